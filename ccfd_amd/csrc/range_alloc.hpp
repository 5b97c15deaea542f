// range_alloc.hpp - first-fit allocator over [0, cap) with coalescing free
// ranges, used for the element blocks, ghosts and boundary-condition
// entries of the batched arena (host code).
#pragma once

#include <iterator>
#include <map>

namespace ccfd {

struct RangeAlloc {
  std::map<int, int> freeR;   // start -> length
  int cap = 0, used = 0;
  void init(int c) {
    cap = c;
    used = 0;
    freeR.clear();
    if (c > 0) freeR[0] = c;
  }
  int alloc(int n) {
    if (n == 0) return 0;
    for (auto it = freeR.begin(); it != freeR.end(); ++it) {
      if (it->second >= n) {
        const int s = it->first, len = it->second;
        freeR.erase(it);
        if (len > n) freeR[s + n] = len - n;
        used += n;
        return s;
      }
    }
    return -1;
  }
  bool can(int n) const {
    if (n == 0) return true;
    for (auto& r : freeR)
      if (r.second >= n) return true;
    return false;
  }
  void release(int s, int n) {
    if (n == 0) return;
    used -= n;
    auto it = freeR.emplace(s, n).first;
    auto nx = std::next(it);
    if (nx != freeR.end() && s + n == nx->first) {
      it->second += nx->second;
      freeR.erase(nx);
    }
    if (it != freeR.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == s) {
        pv->second += it->second;
        freeR.erase(it);
      }
    }
  }
  int high() const {   // end of the highest used position
    if (used == 0) return 0;
    if (freeR.empty()) return cap;
    auto last = std::prev(freeR.end());
    return last->first + last->second == cap ? last->first : cap;
  }
};

}  // namespace ccfd
