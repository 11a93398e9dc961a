/* meshio.c — CPU oracle for the STL and OBJ surface readers.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates src/io/stl.cpp (binary detection by size, little-endian float records, the
 * ASCII "vertex" grammar, build_plc's merge of coincident vertices in first-occurrence
 * order on the exact coordinate triple) and src/io/obj.cpp ("v" and "f" statements,
 * 1-based leading vertex index of each corner), with the shared tokenizer and number
 * parsing of include/cmg/io/detail.hpp (comments from '#', whitespace and commas
 * separate tokens, a number must use its whole token, std::stod / std::stol range
 * errors are failures).
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void seterr(char* err, size_t len, const char* a, const char* b) {
    if (err && len) snprintf(err, len, "%s%s", a, b ? b : "");
}

/* detail.hpp:57-60 (std::stod: strtod, whole token, ERANGE is a failure) */
static int to_real(const char* s, double* out) {
    char* end;
    errno = 0;
    double v = strtod(s, &end);
    if (end == s || *end != '\0' || errno == ERANGE) return 0;
    *out = v;
    return 1;
}

/* detail.hpp:63-66 (std::stol, base 10) */
static int to_long(const char* s, long* out) {
    char* end;
    errno = 0;
    long v = strtol(s, &end, 10);
    if (end == s || *end != '\0' || errno == ERANGE) return 0;
    *out = v;
    return 1;
}

typedef struct {
    char** tok;
    int n, cap;
} toks_t;

/* detail.hpp:24-54: strip '#...' then split on whitespace and commas */
static void tokenize(char* line, toks_t* t) {
    t->n = 0;
    char* h = strchr(line, '#');
    if (h) *h = '\0';
    char* p = line;
    for (;;) {
        while (*p && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' ||
                      *p == '\f' || *p == ','))
            *p++ = '\0';
        if (!*p) break;
        if (t->n == t->cap) {
            t->cap = t->cap ? 2 * t->cap : 16;
            t->tok = (char**)realloc(t->tok, sizeof(char*) * (size_t)t->cap);
        }
        t->tok[t->n++] = p;
        while (*p && !(*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' ||
                       *p == '\f' || *p == ','))
            ++p;
    }
}

typedef struct {
    double* v;
    size_t n, cap;
} dvec;

static void dpush(dvec* d, double x) {
    if (d->n == d->cap) {
        d->cap = d->cap ? 2 * d->cap : 1024;
        d->v = (double*)realloc(d->v, sizeof(double) * d->cap);
    }
    d->v[d->n++] = x;
}

/* Reads a whole line into *buf (growing it); 0 at end of file. */
static int read_line(FILE* f, char** buf, size_t* cap) {
    size_t n = 0;
    int c;
    if (!*buf) {
        *cap = 256;
        *buf = (char*)malloc(*cap);
    }
    while ((c = fgetc(f)) != EOF) {
        if (n + 1 >= *cap) {
            *cap *= 2;
            *buf = (char*)realloc(*buf, *cap);
        }
        if (c == '\n') break;
        (*buf)[n++] = (char)c;
    }
    (*buf)[n] = '\0';
    return !(c == EOF && n == 0);
}

/* ---- merge of coincident vertices (stl.cpp:24-46) -------------------------------- */

typedef struct {
    double x, y, z;
    int64_t i;
} keyed;

static int cmp_keyed(const void* pa, const void* pb) {
    const keyed* a = (const keyed*)pa;
    const keyed* b = (const keyed*)pb;
    /* CoordKey::operator< (x, then y, then z), ties by first occurrence */
    if (a->x != b->x) return a->x < b->x ? -1 : 1;
    if (a->y != b->y) return a->y < b->y ? -1 : 1;
    if (a->z != b->z) return a->z < b->z ? -1 : 1;
    return a->i < b->i ? -1 : (a->i > b->i);
}

/* soup: nv vertices (3 doubles each). Writes pts (first occurrences, in order) and idx
 * (nv entries: the merged index of every vertex); returns the point count. Inputs must
 * hold no NaN (the reference's ordered map has no defined behaviour for them). */
int64_t orc_weld(const double* soup, int64_t nv, double* pts, int32_t* idx) {
    keyed* k = (keyed*)malloc(sizeof(keyed) * (size_t)(nv ? nv : 1));
    for (int64_t i = 0; i < nv; ++i) {
        k[i].x = soup[3 * i];
        k[i].y = soup[3 * i + 1];
        k[i].z = soup[3 * i + 2];
        k[i].i = i;
    }
    qsort(k, (size_t)nv, sizeof(keyed), cmp_keyed);
    int64_t* rep = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv ? nv : 1));
    for (int64_t s = 0; s < nv;) {
        int64_t e = s + 1;
        while (e < nv && k[e].x == k[s].x && k[e].y == k[s].y && k[e].z == k[s].z) ++e;
        for (int64_t j = s; j < e; ++j) rep[k[j].i] = k[s].i; /* first occurrence */
        s = e;
    }
    free(k);
    int64_t* newi = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv ? nv : 1));
    int64_t np = 0;
    for (int64_t i = 0; i < nv; ++i)
        if (rep[i] == i) {
            newi[i] = np;
            pts[3 * np] = soup[3 * i];
            pts[3 * np + 1] = soup[3 * i + 1];
            pts[3 * np + 2] = soup[3 * i + 2];
            ++np;
        }
    for (int64_t i = 0; i < nv; ++i) idx[i] = (int32_t)newi[rep[i]];
    free(newi);
    free(rep);
    return np;
}

/* ---- STL (stl.cpp:94-165) ---------------------------------------------------------- */

/* Reads an STL file into its vertex soup (3 vertices per triangle, doubles). Returns
 * the number of vertices, or -1 with err filled. */
static int64_t stl_soup(const char* path, double** out, int* binary, char* err, size_t len) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        seterr(err, len, "cannot open file for reading: ", path);
        return -1;
    }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    unsigned char* buf = (unsigned char*)malloc((size_t)sz + 1);
    size_t got = fread(buf, 1, (size_t)sz, f);
    fclose(f);
    if ((long)got != sz) {
        free(buf);
        seterr(err, len, "cannot open file for reading: ", path);
        return -1;
    }
    *binary = 0;
    if (sz >= 84) {
        uint32_t n;
        memcpy(&n, buf + 80, 4);
        if ((unsigned long long)sz == 84ull + 50ull * n) {
            *binary = 1;
            double* v = (double*)malloc(sizeof(double) * 9 * (size_t)(n ? n : 1));
            for (uint32_t t = 0; t < n; ++t)
                for (int j = 0; j < 9; ++j) {
                    float x;
                    memcpy(&x, buf + 84 + 50 * (size_t)t + 12 + 4 * j, 4);
                    v[9 * (size_t)t + j] = (double)x;
                }
            free(buf);
            *out = v;
            return 3 * (int64_t)n;
        }
    }
    free(buf);
    f = fopen(path, "r");
    if (!f) {
        seterr(err, len, "cannot open file for reading: ", path);
        return -1;
    }
    dvec d = {0, 0, 0};
    toks_t t = {0, 0, 0};
    char* line = NULL;
    size_t cap = 0;
    int64_t rc = 0;
    while (read_line(f, &line, &cap)) {
        tokenize(line, &t);
        if (t.n == 0) continue;
        if (!strcmp(t.tok[0], "vertex") || !strcmp(t.tok[0], "VERTEX")) {
            if (t.n < 4) {
                seterr(err, len, ".stl: short vertex record", NULL);
                rc = -1;
                break;
            }
            double x, y, z;
            if (!to_real(t.tok[1], &x) || !to_real(t.tok[2], &y) || !to_real(t.tok[3], &z)) {
                seterr(err, len, ".stl: non-numeric vertex coordinate", NULL);
                rc = -1;
                break;
            }
            dpush(&d, x);
            dpush(&d, y);
            dpush(&d, z);
        }
    }
    fclose(f);
    free(line);
    free(t.tok);
    if (rc < 0) {
        free(d.v);
        return -1;
    }
    *out = d.v;
    return (int64_t)(d.n / 3);
}

int64_t orc_read_stl(const char* path, double** pts, int32_t** tris, int64_t* ntris,
                     int* binary, char* err, size_t len) {
    *pts = NULL;
    *tris = NULL;
    *ntris = 0;
    double* soup = NULL;
    int64_t nv = stl_soup(path, &soup, binary, err, len);
    if (nv < 0) return -1;
    /* stl.cpp:50-51 */
    if (nv == 0 || nv % 3 != 0) {
        free(soup);
        seterr(err, len, ".stl: vertex count is not a positive multiple of 3", NULL);
        return -1;
    }
    *pts = (double*)malloc(sizeof(double) * 3 * (size_t)nv);
    *tris = (int32_t*)malloc(sizeof(int32_t) * (size_t)nv);
    int64_t np = orc_weld(soup, nv, *pts, *tris);
    free(soup);
    *ntris = nv / 3;
    return np;
}

/* Raw soup of an STL file (no merge): 3 doubles per vertex; returns vertices or -1. */
int64_t orc_stl_soup(const char* path, double** soup, int* binary, char* err, size_t len) {
    return stl_soup(path, soup, binary, err, len);
}

/* ---- OBJ (obj.cpp:33-72) ------------------------------------------------------------ */

int orc_read_obj(const char* path, double** pts, int64_t* npts, int32_t** off,
                 int32_t** idx, int64_t* nfacets, char* err, size_t len) {
    *pts = NULL;
    *off = NULL;
    *idx = NULL;
    *npts = *nfacets = 0;
    FILE* f = fopen(path, "r");
    if (!f) {
        seterr(err, len, "cannot open file for reading: ", path);
        return -1;
    }
    dvec P = {0, 0, 0};
    size_t noff = 1, offcap = 1024, nidx = 0, idxcap = 1024;
    int32_t* O = (int32_t*)malloc(sizeof(int32_t) * offcap);
    int32_t* I = (int32_t*)malloc(sizeof(int32_t) * idxcap);
    O[0] = 0;
    toks_t t = {0, 0, 0};
    char* line = NULL;
    size_t cap = 0;
    int rc = 0;
    while (read_line(f, &line, &cap)) {
        tokenize(line, &t);
        if (t.n == 0) continue;
        if (!strcmp(t.tok[0], "v")) {
            if (t.n < 4) {
                seterr(err, len, ".obj: short vertex record", NULL);
                rc = -1;
                break;
            }
            double x, y, z;
            if (!to_real(t.tok[1], &x) || !to_real(t.tok[2], &y) || !to_real(t.tok[3], &z)) {
                seterr(err, len, ".obj: non-numeric vertex coordinate", NULL);
                rc = -1;
                break;
            }
            dpush(&P, x);
            dpush(&P, y);
            dpush(&P, z);
        } else if (!strcmp(t.tok[0], "f")) {
            if (t.n < 4) {
                seterr(err, len, ".obj: face needs at least 3 corners", NULL);
                rc = -1;
                break;
            }
            for (int k = 1; k < t.n; ++k) {
                char* slash = strchr(t.tok[k], '/');
                if (slash) *slash = '\0';
                long v = 0;
                if (!to_long(t.tok[k], &v)) {
                    if (slash) *slash = '/';
                    seterr(err, len, ".obj: malformed face corner: ", t.tok[k]);
                    rc = -1;
                    break;
                }
                if (nidx == idxcap) {
                    idxcap *= 2;
                    I = (int32_t*)realloc(I, sizeof(int32_t) * idxcap);
                }
                I[nidx++] = (int32_t)(v - 1);
            }
            if (rc) break;
            if (noff == offcap) {
                offcap *= 2;
                O = (int32_t*)realloc(O, sizeof(int32_t) * offcap);
            }
            O[noff++] = (int32_t)nidx;
        }
    }
    fclose(f);
    free(line);
    free(t.tok);
    if (rc) {
        free(P.v);
        free(O);
        free(I);
        return -1;
    }
    *pts = P.v;
    *npts = (int64_t)(P.n / 3);
    *off = O;
    *idx = I;
    *nfacets = (int64_t)(noff - 1);
    return 0;
}

void orc_free(void* p) { free(p); }
