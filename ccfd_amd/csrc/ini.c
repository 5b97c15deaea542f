/*
 * ini.c - parameter files with the reference's rules (readInTools,
 * src/readInTools.c:38-405): keys are case-insensitive, '!' and '#' start
 * comments, values keep only [A-Za-z0-9_.,-] so "(/1.0, 0.0/)" reads as
 * "1.0,0.0", and every key is consumed when read, which makes repeated keys
 * (BCtype, rho, meshBCtype, ...) positional. Integers parse through strtod.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccfd_dropin.h"

#define KLEN 256

typedef struct ini_cmd {
	char key[KLEN];
	char val[KLEN];
	int used;
} ini_cmd;

struct ccfd_ini {
	ini_cmd *cmd;
	int n;
};

ccfd_ini *ccfd_ini_read(const char *path, char *err)
{
	FILE *f = fopen(path, "r");
	if (!f) {
		if (err)
			snprintf(err, 256, "could not open parameter file '%s'", path);
		return NULL;
	}
	ccfd_ini *ini = calloc(1, sizeof(*ini));
	int cap = 64;
	ini->cmd = calloc(cap, sizeof(ini_cmd));
	char line[2 * KLEN];
	while (fgets(line, sizeof(line), f)) {
		char key[KLEN], val[KLEN];
		size_t i, j = 0, L = strlen(line);
		for (i = 0; i < L; ++i) {
			unsigned char ch = (unsigned char)line[i];
			if (isalnum(ch) || ch == '_') {
				if (j < KLEN - 1)
					key[j++] = (char)tolower(ch);
			} else if (ch == '!' || ch == '#' || ch == '=') {
				break;
			}
		}
		key[j] = '\0';
		if (j == 0)
			continue;
		j = 0;
		for (; i < L; ++i) {
			unsigned char ch = (unsigned char)line[i];
			if (isalnum(ch) || ch == '_' || ch == '.' || ch == ',' ||
					ch == '-') {
				if (j < KLEN - 1)
					val[j++] = (char)ch;
			} else if (ch == '!' || ch == '#') {
				break;
			}
		}
		val[j] = '\0';
		if (ini->n == cap) {
			cap *= 2;
			ini->cmd = realloc(ini->cmd, cap * sizeof(ini_cmd));
		}
		strcpy(ini->cmd[ini->n].key, key);
		strcpy(ini->cmd[ini->n].val, val);
		ini->cmd[ini->n].used = 0;
		ini->n++;
	}
	fclose(f);
	return ini;
}

void ccfd_ini_free(ccfd_ini *ini)
{
	if (!ini)
		return;
	free(ini->cmd);
	free(ini);
}

static void lower(const char *in, char *out)
{
	size_t i;
	for (i = 0; in[i] && i < KLEN - 1; ++i)
		out[i] = (char)tolower((unsigned char)in[i]);
	out[i] = '\0';
}

/* first unconsumed value of key, or def; NULL and *missing if neither */
static const char *take(ccfd_ini *ini, const char *key, const char *def,
		int *missing)
{
	char k[KLEN];
	lower(key, k);
	*missing = 0;
	for (int i = 0; i < ini->n; ++i) {
		if (!ini->cmd[i].used && !strcmp(ini->cmd[i].key, k)) {
			ini->cmd[i].used = 1;
			return ini->cmd[i].val;
		}
	}
	if (!def)
		*missing = 1;
	return def;
}

int ccfd_ini_count(ccfd_ini *ini, const char *key)
{
	char k[KLEN];
	lower(key, k);
	int c = 0;
	for (int i = 0; i < ini->n; ++i)
		if (!ini->cmd[i].used && !strcmp(ini->cmd[i].key, k))
			c++;
	return c;
}

int ccfd_ini_str(ccfd_ini *ini, const char *key, const char *def, char *out,
		size_t n)
{
	int miss;
	const char *v = take(ini, key, def, &miss);
	if (miss)
		return -1;
	snprintf(out, n, "%s", v);
	return 0;
}

int ccfd_ini_dbl(ccfd_ini *ini, const char *key, const char *def, double *out)
{
	int miss;
	const char *v = take(ini, key, def, &miss);
	if (miss)
		return -1;
	*out = strtod(v, NULL);
	return 0;
}

int ccfd_ini_long(ccfd_ini *ini, const char *key, const char *def, long *out)
{
	double d;
	if (ccfd_ini_dbl(ini, key, def, &d))
		return -1;
	*out = (long)d;
	return 0;
}

int ccfd_ini_bool(ccfd_ini *ini, const char *key, const char *def, int *out)
{
	int miss;
	const char *v = take(ini, key, def, &miss);
	if (miss)
		return -1;
	*out = !strcmp(v, "T") || !strcmp(v, "True") || !strcmp(v, "t") ||
		!strcmp(v, "true");
	return 0;
}

int ccfd_ini_dbl_array(ccfd_ini *ini, const char *key, int n, const char *def,
		double *out)
{
	int miss;
	const char *v = take(ini, key, def, &miss);
	if (miss)
		return -1;
	char buf[KLEN];
	snprintf(buf, sizeof(buf), "%s", v);
	char *save = NULL, *tok = strtok_r(buf, ",", &save);
	for (int i = 0; i < n; ++i) {
		out[i] = tok ? strtod(tok, NULL) : 0.0;
		tok = strtok_r(NULL, ",", &save);
	}
	return 0;
}
