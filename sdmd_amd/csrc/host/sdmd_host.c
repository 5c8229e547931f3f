/*
 * sdmd_host.c - setup of one sDMD system from the reference's input formats.
 *
 * Every routine here restates the behaviour of the reference's setup code,
 * cited per function, including its order-dependent details (prepend order
 * of bond/constraint lists, the one-directional connection-map updates, the
 * 6-decimal text round trip of coordinates, the RNG draw order), because the
 * event loop downstream must start from a bit-identical state.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sdmd_host.h"
#include "sdmd_host_internal.h"

/* ------------------------------------------------------------------ */
/* glibc TYPE_3 additive-feedback generator (the rand() the reference   */
/* calls, ToolFunctions.c:519): r[i] = r[i-31] + r[i-3] mod 2^32,       */
/* output r >> 1; seeded by a Lehmer LCG (16807, 2^31-1) computed with  */
/* Schrage's method, then 310 outputs discarded.                        */
/* ------------------------------------------------------------------ */
void sdmd_rand_seed(sdmd_rand_state *st, uint32_t seed) {
    int32_t *r = st->r;
    if (seed == 0) seed = 1;
    int32_t word = (int32_t)seed;
    r[0] = word;
    for (int i = 1; i < 31; i++) {
        long hi = word / 127773;
        long lo = word % 127773;
        long w = 16807 * lo - 2836 * hi;
        if (w < 0) w += 2147483647;
        word = (int32_t)w;
        r[i] = word;
    }
    st->f = 3;
    st->b = 0;
    st->calls = 0;
    for (int k = 0; k < 310; k++) sdmd_rand_next(st);
    st->calls = 0;
}

int32_t sdmd_rand_next(sdmd_rand_state *st) {
    uint32_t v = (uint32_t)st->r[st->f] + (uint32_t)st->r[st->b];
    st->r[st->f] = (int32_t)v;
    int32_t out = (int32_t)(v >> 1);
    if (++st->f >= 31) st->f = 0;
    if (++st->b >= 31) st->b = 0;
    st->calls++;
    return out;
}

/* RandomValue(0,1), ToolFunctions.c:517-522 */
static double rand01(sdmd_rand_state *st) {
    return 0 + (double)sdmd_rand_next(st) / ((double)2147483647 / (1 - 0));
}

/* Maxwell_Boltzmann_Distribution(0, 1) output value, ToolFunctions.c:544-569 */
static double mb_normal(sdmd_rand_state *st) {
    double u0, u1, s, choose, out;
    for (;;) {
        do {
            u0 = 2 * rand01(st) - 1;
            u1 = 2 * rand01(st) - 1;
            s = u0 * u0 + u1 * u1;
        } while (s >= 1);
        s = sqrt(-2 * log(s) / s);
        choose = rand01(st);
        out = (choose >= 0.5) ? u0 * s : u1 * s;
        if (out != 0) return out;
    }
}

/* ------------------------------------------------------------------ */
/* setup-time data model                                              */
/* ------------------------------------------------------------------ */

static int fail(sdmd_host_system *s, int code, const char *fmt, ...) {
    if (s && s->err && s->errlen > 0) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(s->err, (size_t)s->errlen, fmt, ap);
        va_end(ap);
    }
    return code;
}

static void *xrealloc(void *p, size_t n) {
    void *q = realloc(p, n ? n : 1);
    if (!q) { fprintf(stderr, "sdmd_host: out of memory\n"); abort(); }
    return q;
}

static int new_steplist(sdmd_host_system *s) {
    if (s->nsteplist == s->capsteplist) {
        s->capsteplist = s->capsteplist ? 2 * s->capsteplist : 256;
        s->steplists = xrealloc(s->steplists, sizeof(steplist_t) * s->capsteplist);
    }
    memset(&s->steplists[s->nsteplist], 0, sizeof(steplist_t));
    return s->nsteplist++;
}

static void steplist_push(steplist_t *l, double d, double e) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 8;
        l->v = xrealloc(l->v, sizeof(step_t) * l->cap);
    }
    l->v[l->n].d = d;
    l->v[l->n].e = e;
    l->v[l->n].acc = 0;
    l->n++;
}

/* CalAccumulatedPoten, ToolFunctions.c:897-919 */
static void accumulate(steplist_t *l) {
    double acc = 0;
    for (int k = 0; k < l->n; k++) acc += l->v[k].e;
    for (int k = 0; k < l->n; k++) { l->v[k].acc = acc; acc -= l->v[k].e; }
}

static void link_prepend(linklist_t *l, link_t e) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 8;
        l->v = xrealloc(l->v, sizeof(link_t) * l->cap);
    }
    memmove(l->v + 1, l->v, sizeof(link_t) * l->n);
    l->v[0] = e;
    l->n++;
}

static void link_remove_first(linklist_t *l, int j) {   /* RemoveConstr, Initialization.c:1233-1259 */
    for (int k = 0; k < l->n; k++)
        if (l->v[k].j == j) {
            memmove(l->v + k, l->v + k + 1, sizeof(link_t) * (l->n - k - 1));
            l->n--;
            return;
        }
}

static int link_find(const linklist_t *l, int j) {
    for (int k = 0; k < l->n; k++)
        if (l->v[k].j == j) return k;
    return -1;
}

static int *cm_ref(cmrow_t *row, int j) {
    for (int k = 0; k < row->n; k++)
        if (row->v[k].j == j) return &row->v[k].bits;
    if (row->n == row->cap) {
        row->cap = row->cap ? 2 * row->cap : 8;
        row->v = xrealloc(row->v, sizeof(cm_t) * row->cap);
    }
    row->v[row->n].j = j;
    row->v[row->n].bits = 0;
    return &row->v[row->n++].bits;
}

static int cm_get(const cmrow_t *row, int j) {
    for (int k = 0; k < row->n; k++)
        if (row->v[k].j == j) return row->v[k].bits;
    return 0;
}

/* ------------------------------------------------------------------ */
/* type maps, Models.c                                                */
/* ------------------------------------------------------------------ */
static const char *AA_NAMES[] = {"", "ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS",
                                 "ILE", "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP",
                                 "TYR", "VAL", "SOL", "GEL", "PHO", "PHI", "RPH"};

static int aa_model(const char *n) {          /* AAModel, Models.c:65-126 */
    for (int k = 1; k <= 25; k++)
        if (strcmp(n, AA_NAMES[k]) == 0) return k;
    return -1;
}

static int atom_model(const char *t) {        /* AtomModel, Models.c:129-209 */
    static const struct { const char *n; int v; } m[] = {
        {"OZB", 26}, {"NZB", 18}, {"CRB", 6}, {"CA", 9}, {"CA1", 10}, {"CA2", 11}, {"CA3", 12},
        {"CA1P", 14}, {"CA2P", 15}, {"CM1", 13}, {"CM1P", 16}, {"CR", 5}, {"CRN", 7},
        {"CRNP", 8}, {"NCK", 21}, {"NCR", 22}, {"NR", 23}, {"NZNP", 19}, {"NZNQ", 20},
        {"NZ", 17}, {"OC", 28}, {"OW", 29}, {"OZH", 25}, {"OZ", 24}, {"SGNP", 30}, {"SG", 31},
        {"OZB_HB", 27}, {"HB", 2}, {"HO", 1}, {"HN", 1}, {"HS", 1}, {"HW", 1}, {"H", 1},
        {"H_HB", 3}, {"HB_HB", 4}, {"GEL", 32}};
    for (size_t k = 0; k < sizeof m / sizeof m[0]; k++)
        if (strcmp(t, m[k].n) == 0) return m[k].v;
    return -1;
}

/* ------------------------------------------------------------------ */
/* a tiny scanf-like reader over a whole file                         */
/* ------------------------------------------------------------------ */
typedef struct { char *buf; size_t n, pos; } rd_t;

static int rd_open(rd_t *r, const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    r->buf = malloc((size_t)n + 1);
    r->n = fread(r->buf, 1, (size_t)n, f);
    r->buf[r->n] = 0;
    r->pos = 0;
    fclose(f);
    return 0;
}
static void rd_close(rd_t *r) { free(r->buf); r->buf = NULL; }

/* fscanf("%s"): skip white space, read a token; 0 at end of input */
static int rd_tok(rd_t *r, char *out, size_t cap) {
    while (r->pos < r->n && isspace((unsigned char)r->buf[r->pos])) r->pos++;
    if (r->pos >= r->n) return 0;
    size_t k = 0;
    while (r->pos < r->n && !isspace((unsigned char)r->buf[r->pos])) {
        if (k + 1 < cap) out[k++] = r->buf[r->pos];
        r->pos++;
    }
    out[k] = 0;
    return 1;
}

/* fgets: the rest of the current line including its newline */
static void rd_line(rd_t *r, char *out, size_t cap) {
    size_t k = 0;
    while (r->pos < r->n) {
        char c = r->buf[r->pos++];
        if (k + 1 < cap) out[k++] = c;
        if (c == '\n') break;
    }
    out[k] = 0;
}

static int rd_double(rd_t *r, double *v) {
    char t[128];
    if (!rd_tok(r, t, sizeof t)) return 0;
    char *end;
    *v = strtod(t, &end);
    return end != t;
}

/* ------------------------------------------------------------------ */
/* ReadParameter, Initialization.c:512-730                            */
/* ------------------------------------------------------------------ */
static int read_parameter(sdmd_host_system *s) {
    char path[1200], key[256], val[256], line[1024];
    snprintf(path, sizeof path, "%s/parameter.txt", s->datadir);
    FILE *f = fopen(path, "r");
    if (!f) return fail(s, SDMD_EIO, "cannot open %s", path);
    char *fields[16];
    char vals[16][256];
    int n = 0;
    while (n < 16 && fgets(line, sizeof line, f)) {
        if (sscanf(line, "%255s%255s", key, val) < 2) continue;
        snprintf(vals[n], sizeof vals[n], "%s", val);
        fields[n] = vals[n];
        n++;
    }
    fclose(f);
    if (n < 16) return fail(s, SDMD_EFORMAT, "parameter.txt: expected 16 lines, got %d", n);
    snprintf(s->neworcontinue, sizeof s->neworcontinue, "%s", fields[0]);
    s->timestep = atof(fields[1]);
    snprintf(s->temptype, sizeof s->temptype, "%s", fields[2]);
    if (strcmp(s->temptype, "no")) s->T = atof(s->temptype);
    s->outputrate = atof(fields[3]);
    s->cutoffr = atof(fields[4]);
    snprintf(s->method, sizeof s->method, "%s", fields[5]);
    snprintf(s->thermostat, sizeof s->thermostat, "%s", fields[6]);
    s->thermoF = atof(fields[7]);
    snprintf(s->wallexist, sizeof s->wallexist, "%s", fields[9]);
    snprintf(s->walltype, sizeof s->walltype, "%s", fields[10]);
    if (strncmp(s->neworcontinue, "new", 1) && strncmp(s->neworcontinue, "continue", 1))
        return fail(s, SDMD_EFORMAT, "Continue? must be 'new' or 'continue' (got %s)", s->neworcontinue);
    if (strncmp(s->method, "Ding", 1))
        return fail(s, SDMD_EUNSUPPORTED, "only the Ding force field is supported");
    if (strcmp(fields[8], "no"))
        return fail(s, SDMD_EUNSUPPORTED, "CGModel %s is outside the supported path", fields[8]);
    for (int k = 11; k < 16; k++)
        if (strcmp(fields[k], "no"))
            return fail(s, SDMD_EUNSUPPORTED, "parameter line %d (%s) must be 'no'", k + 1, fields[k]);
    if (strcmp(s->thermostat, "no") && strncmp(s->thermostat, "Andersen", 1))
        return fail(s, SDMD_EUNSUPPORTED, "thermostat %s", s->thermostat);
    if (strcmp(s->wallexist, "no")) {
        if (strcmp(s->wallexist, "smooth") || strncmp(s->walltype, "cylinder", 1))
            return fail(s, SDMD_EUNSUPPORTED, "only a smooth cylinder wall is supported");
    }
    return SDMD_OK;
}

/* ------------------------------------------------------------------ */
/* residue templates                                                  */
/* ------------------------------------------------------------------ */
static int template_path(sdmd_host_system *s, const char *aa, char *out, size_t cap) {
    snprintf(out, cap, "%s/Library_%s/AA/%s.txt", s->datadir, s->method, aa);
    return 0;
}

/* ATOM names of a template, in order (ReadGROFile 924-928) */
static int template_atoms(sdmd_host_system *s, const char *aa, char list[][8], int *num) {
    char path[1300], tok[256], rest[1024];
    template_path(s, strcmp(aa, "GLY") == 0 ? "ALA" : aa, path, sizeof path);
    rd_t r;
    if (rd_open(&r, path)) return fail(s, SDMD_EIO, "cannot open template %s", path);
    rd_line(&r, rest, sizeof rest);
    *num = 0;
    while (rd_tok(&r, tok, sizeof tok) && strcmp(tok, "ATOM") == 0) {
        if (*num >= 20) { rd_close(&r); return fail(s, SDMD_EFORMAT, "template %s has >20 atoms", aa); }
        rd_tok(&r, list[*num], 8);
        rd_line(&r, rest, sizeof rest);
        (*num)++;
    }
    rd_close(&r);
    return SDMD_OK;
}

/* CheckValidAtom, Initialization.c:1342-1357 */
static int check_valid_atom(char *name, char list[][8], int num) {
    for (int i = 0; i < num; i++) {
        if (strcmp(name, list[i]) == 0) { strcpy(list[i], "0"); return 1; }
        if (strcmp(name, "O1") == 0) { strcpy(list[i], "0"); strcpy(name, "O"); return 1; }
    }
    return 0;
}

/* RemoveSpace, Initialization.c:1360-1372 */
static void remove_space(char *str) {
    int n = 0;
    char *p = str;
    while (*p == ' ') p++;
    while (*p != ' ' && *p) str[n++] = *p++;
    str[n] = 0;
}

/* text round trip of ReadGROFile 937-941 + ReadCoordinate 1027-1030 */
static double nm_to_angstrom_6dp(double nm) {
    char b[64];
    snprintf(b, sizeof b, "%16.6lf", nm * 10);
    return strtod(b, NULL);
}

/* ReadGROFile (Initialization.c:877-989) + ReadCoordinate (992-1057) */
static int read_gro(sdmd_host_system *s, const char *gro) {
    char path[1300], line[1024], resn[8], atn[8], tmpAA[256] = {0}, tok[256];
    char list[20][8];
    int num = 0, tmpInt = 0;
    snprintf(path, sizeof path, "%s/%s", s->datadir, gro);
    rd_t r;
    if (rd_open(&r, path)) return fail(s, SDMD_EIO, "cannot open %s", path);
    rd_line(&r, line, sizeof line);
    if (!rd_tok(&r, tok, sizeof tok)) { rd_close(&r); return fail(s, SDMD_EFORMAT, "gro: no atom count"); }
    int count = atoi(tok);
    double *cx = malloc(sizeof(double) * 3 * (count + 1));
    char (*names)[8] = calloc((size_t)count + 1, 8);
    int natom = 0;
    s->naa = 0;
    s->nprot = 0;
    s->aa = xrealloc(NULL, sizeof(aa_t));
    s->pep = xrealloc(NULL, sizeof(pep_t));
    memset(s->aa, 0, sizeof(aa_t));
    memset(s->pep, 0, sizeof(pep_t));
    for (int i = 1; i <= count; i++) {
        while (r.pos < r.n && isspace((unsigned char)r.buf[r.pos])) r.pos++;
        char *end;
        long rInt = strtol(r.buf + r.pos, &end, 0);
        r.pos = (size_t)(end - r.buf);
        memset(resn, 0, sizeof resn);
        memset(atn, 0, sizeof atn);
        for (int k = 0; k < 5 && r.pos < r.n; k++) resn[k] = r.buf[r.pos++];
        for (int k = 0; k < 5 && r.pos < r.n; k++) atn[k] = r.buf[r.pos++];
        rd_tok(&r, tok, sizeof tok);
        remove_space(resn);
        remove_space(atn);
        int newres = (tmpInt != rInt || strcmp(resn, tmpAA));
        if (newres) {
            int rc = template_atoms(s, resn, list, &num);
            if (rc) { rd_close(&r); free(cx); free(names); return rc; }
        }
        rd_line(&r, line, sizeof line);
        if (check_valid_atom(atn, list, num)) {
            natom++;
            double v[3];
            if (sscanf(line, "%lf%lf%lf", &v[0], &v[1], &v[2]) != 3) {
                rd_close(&r); free(cx); free(names);
                return fail(s, SDMD_EFORMAT, "gro: bad coordinates on atom %d", i);
            }
            for (int k = 0; k < 3; k++) cx[3 * natom + k] = nm_to_angstrom_6dp(v[k]);
            snprintf(names[natom], 8, "%s", atn);
        }
        if (newres) {
            s->naa++;
            s->aa = xrealloc(s->aa, sizeof(aa_t) * (s->naa + 1));
            aa_t *a = &s->aa[s->naa];
            memset(a, 0, sizeof *a);
            snprintf(a->name, sizeof a->name, "%s", resn);
            int newprot = (rInt < tmpInt || s->nprot == 0);
            a->protein = newprot ? s->nprot + 1 : s->nprot;
            s->aa[s->naa - 1].end = natom - 1;
            a->start = natom;
            a->type = aa_model(a->name);
            if (a->type < 0) { rd_close(&r); free(cx); free(names); return fail(s, SDMD_EFORMAT, "unknown residue %s", resn); }
            if (newprot) {
                s->nprot++;
                s->pep = xrealloc(s->pep, sizeof(pep_t) * (s->nprot + 1));
                memset(&s->pep[s->nprot], 0, sizeof(pep_t));
                s->pep[s->nprot - 1].endAA = s->naa - 1;
                s->pep[s->nprot].startAA = s->naa;
                s->pep[s->nprot - 1].endAtom = natom - 1;
                s->pep[s->nprot].startAtom = natom;
            }
            tmpInt = (int)rInt;
            snprintf(tmpAA, sizeof tmpAA, "%s", resn);
        }
    }
    double b[3];
    for (int k = 0; k < 3; k++) {
        if (!rd_double(&r, &b[k])) { rd_close(&r); free(cx); free(names); return fail(s, SDMD_EFORMAT, "gro: no box line"); }
        s->box[k + 1] = b[k] * 10;
    }
    rd_close(&r);
    s->aa[s->naa].end = natom;
    s->pep[s->nprot].endAA = s->naa;
    s->pep[s->nprot].endAtom = natom;

    /* ReadCoordinate: atoms, sequence numbering, residue types */
    s->natom = natom;
    int N1 = natom + 1;
    s->p = calloc((size_t)N1, sizeof(atomprop_t));
    s->x = calloc((size_t)N1, sizeof(double)); s->y = calloc((size_t)N1, sizeof(double));
    s->z = calloc((size_t)N1, sizeof(double)); s->vx = calloc((size_t)N1, sizeof(double));
    s->vy = calloc((size_t)N1, sizeof(double)); s->vz = calloc((size_t)N1, sizeof(double));
    s->cx = calloc((size_t)N1, sizeof(int)); s->cy = calloc((size_t)N1, sizeof(int));
    s->cz = calloc((size_t)N1, sizeof(int));
    s->role = calloc((size_t)N1, 1);
    s->hbnb = calloc((size_t)N1, sizeof(int));
    s->bond = calloc((size_t)N1, sizeof(linklist_t));
    s->constr = calloc((size_t)N1, sizeof(linklist_t));
    s->cmap = calloc((size_t)N1, sizeof(cmrow_t));
    int AANum = 1, pepNum = 1;
    for (int i = 1; i <= natom; i++) {
        snprintf(s->p[i].name, 8, "%s", names[i]);
        s->x[i] = cx[3 * i]; s->y[i] = cx[3 * i + 1]; s->z[i] = cx[3 * i + 2];
        if (i > s->aa[AANum].end) AANum++;
        if (i > s->pep[pepNum].endAtom) pepNum++;
        s->p[i].aanum = AANum - s->pep[pepNum - 1].endAA;
        s->p[i].protein = pepNum;
        snprintf(s->p[i].aa, 8, "%s", s->aa[AANum].name);
        s->p[i].aatype = aa_model(s->p[i].aa);
    }
    free(cx);
    free(names);
    return SDMD_OK;
}

/* CreateCell, Initialization.c:1060-1073 */
static void create_cell(sdmd_host_system *s) {
    for (int i = 1; i <= 3; i++) {
        s->cellnum[i] = (int)floor(s->box[i] / s->cutoffr);
        if (s->cellnum[i] < 3) s->cellnum[i] = 3;
    }
    for (int i = 1; i <= 3; i++) s->cellsize[i] = s->box[i] / s->cellnum[i];
}

/* FindTargetAtom, Initialization.c:1375-1397 */
static int find_target(sdmd_host_system *s, int AANum, const char *t) {
    int start, end;
    if (t[0] == '+') {
        if (AANum == s->naa || s->aa[AANum + 1].protein != s->aa[AANum].protein) return -1;
        start = s->aa[AANum + 1].start;
        end = s->aa[AANum + 1].end;
    } else {
        start = s->aa[AANum].start;
        end = s->aa[AANum].end;
    }
    const char *name = t[0] == '+' ? t + 1 : t;
    for (int i = start; i <= end; i++)
        if (i >= 0 && i <= s->natom && strcmp(s->p[i].name, name) == 0) return i;
    return 0;
}

static const char *aa_name(sdmd_host_system *s, int k) {
    return (k >= 1 && k <= s->naa) ? s->aa[k].name : "";
}

/* FindNextBondList, Initialization.c:1262-1296 */
static int next_bond_list(sdmd_host_system *s, int num, int target, double dmin, double dmax, int type) {
    linklist_t *l = &s->bond[num];
    int k = link_find(l, target);
    if (k >= 0) {
        if (l->v[k].dmin != dmin || l->v[k].dmax != dmax) {
            if (strcmp(s->p[num].aa, "PRO") && strcmp(s->p[target].aa, "PRO")) return 0;
            l->v[k].dmin = dmin;
            l->v[k].dmax = dmax;
        }
        return 0;
    }
    link_t e = {target, dmin, dmax, -1};
    link_prepend(l, e);
    if (type == 2) link_remove_first(&s->constr[num], target);
    return 1;
}

/* ConstrList, Initialization.c:1299-1339 */
static int constr_list(sdmd_host_system *s, int num, int target, int nstep, rd_t *r, int type) {
    linklist_t *l = &s->constr[num];
    if (link_find(l, target) >= 0) return 0;
    link_t e = {target, 0, 0, -1};
    rd_double(r, &e.dmin);
    e.dmin *= e.dmin;
    if (nstep > 0) {
        e.steps = new_steplist(s);
        for (int i = 0; i < nstep; i++) {
            double d = 0, en = 0;
            rd_double(r, &d);
            rd_double(r, &en);
            d *= d;
            steplist_push(&s->steplists[e.steps], d, en);
        }
    }
    rd_double(r, &e.dmax);
    e.dmax *= e.dmax;
    link_prepend(l, e);
    if (type == 2) link_remove_first(&s->bond[num], target);
    return 1;
}

/* BondDuplicate, Initialization.c:1400-1421 */
static void bond_duplicate(sdmd_host_system *s, int AANum) {
    for (int t = s->aa[AANum].start; t <= s->aa[AANum].end; t++) {
        int from = (s->p[t].aanum == 1) ? s->aa[AANum].start : s->aa[AANum - 1].start;
        for (int i = from; i <= s->aa[AANum].end; i++) {
            linklist_t *l = &s->bond[i];
            for (int k = 0; k < l->n; k++) {
                if (l->v[k].j == t) {
                    int flag = next_bond_list(s, t, i, l->v[k].dmin, l->v[k].dmax, 0);
                    if (flag == 1) *cm_ref(&s->cmap[t], i) |= SDMD_BOND_CONNECT;
                    break;
                }
            }
        }
    }
}

/* ConstrDuplicate, Initialization.c:1424-1465 */
static void constr_duplicate(sdmd_host_system *s, int AANum) {
    for (int t = s->aa[AANum].start; t <= s->aa[AANum].end; t++) {
        int from = (s->p[t].aanum == 1) ? s->aa[AANum].start : s->aa[AANum - 1].start;
        for (int i = from; i <= s->aa[AANum].end; i++) {
            linklist_t *l = &s->constr[i];
            for (int k = 0; k < l->n; k++) {
                if (l->v[k].j == t) {
                    if (link_find(&s->constr[t], i) >= 0) continue;
                    link_t e = {i, l->v[k].dmin, l->v[k].dmax, l->v[k].steps};
                    link_prepend(&s->constr[t], e);
                    *cm_ref(&s->cmap[t], i) |= SDMD_CONSTRAINT_CONNECT;
                    break;
                }
            }
        }
    }
}

/* ScanAA, Initialization.c:1098-1230 */
static int scan_aa(sdmd_host_system *s, int AANum, int type) {
    char aaname[32], path[1300], tok[256], rest[1024], n1[64], n2[64];
    if (type == 2) snprintf(aaname, sizeof aaname, "PRO_INTER");
    else if (type == 1) snprintf(aaname, sizeof aaname, "ALA");
    else snprintf(aaname, sizeof aaname, "%s", s->aa[AANum].name);
    template_path(s, aaname, path, sizeof path);
    rd_t r;
    if (rd_open(&r, path)) return fail(s, SDMD_EIO, "cannot open template %s", path);
    rd_line(&r, rest, sizeof rest);
    const char *thisAA = s->aa[AANum].name;
    const char *nextAA = aa_name(s, AANum + 1);
    while (rd_tok(&r, tok, sizeof tok)) {
        if (strcmp(tok, "ATOM") == 0) {
            if (type == 0 && !s->cont) {    /* a continue run keeps the saved types and masses */
                char tname[64], e0[64], e1[64];
                double mass = 0;
                rd_tok(&r, n1, sizeof n1);
                int num = find_target(s, AANum, n1);
                rd_tok(&r, tname, sizeof tname);
                rd_double(&r, &mass);
                rd_tok(&r, e0, sizeof e0);
                rd_tok(&r, e1, sizeof e1);
                if (num > 0) {
                    s->p[num].mass = mass;
                    snprintf(s->p[num].extra0, 16, "%s", e0);
                    snprintf(s->p[num].extra1, 16, "%s", e1);
                    s->p[num].type = atom_model(tname);
                    if (s->p[num].type < 0) { rd_close(&r); return fail(s, SDMD_EFORMAT, "unknown atom type %s", tname); }
                }
            } else {
                rd_line(&r, rest, sizeof rest);
            }
        } else if (!strcmp(tok, "BOND") || !strcmp(tok, "ANGLE") || !strcmp(tok, "INTER")) {
            rd_tok(&r, n1, sizeof n1);
            int num = find_target(s, AANum, n1);
            if (num > 0) {
                double length = 0, delta = 0;
                rd_tok(&r, n2, sizeof n2);
                rd_double(&r, &length);
                rd_double(&r, &delta);
                int t2 = find_target(s, AANum, n2);
                if (t2 == 0) {
                    if (strncmp(n2, "H", 1) == 0 || (strcmp(nextAA, "GLY") == 0 && strcmp(n2, "+CB") == 0) ||
                        (strcmp(thisAA, "GLY") == 0 && strcmp(n2, "CB") == 0) ||
                        (strcmp(nextAA, "PRO") == 0 && strcmp(n2, "+H") == 0))
                        continue;
                    rd_close(&r);
                    return fail(s, SDMD_EFORMAT, "residue %d: atom %s not found", AANum, n2);
                } else if (t2 == -1) {
                    continue;
                }
                double dmin = length * (1 - delta); dmin *= dmin;
                double dmax = length * (1 + delta); dmax *= dmax;
                int flag = next_bond_list(s, num, t2, dmin, dmax, type);
                if (flag == 1) {
                    int *b = cm_ref(&s->cmap[num], t2);
                    if (*b & SDMD_CONSTRAINT_CONNECT) *b ^= SDMD_CONSTRAINT_CONNECT;
                    *b |= SDMD_BOND_CONNECT;
                }
            } else if (num == 0) {
                if (strncmp(n1, "H", 1) == 0) { rd_line(&r, rest, sizeof rest); continue; }
                if (strcmp(n1, "CB") == 0 && strcmp(thisAA, "GLY") == 0) { rd_line(&r, rest, sizeof rest); continue; }
                rd_close(&r);
                return fail(s, SDMD_EFORMAT, "residue %d: atom %s not found", AANum, n1);
            }
        } else if (strncmp(tok, "CONSTR", 6) == 0) {
            int nstep = tok[6] - '0';
            rd_tok(&r, n1, sizeof n1);
            rd_tok(&r, n2, sizeof n2);
            int num = find_target(s, AANum, n1);
            if (num <= 0 && !(strcmp(thisAA, "GLY") == 0 && strcmp(n1, "CB") == 0)) {
                rd_close(&r);
                return fail(s, SDMD_EFORMAT, "residue %d: atom %s not found", AANum, n1);
            }
            int t2 = find_target(s, AANum, n2);
            if (t2 > 0 && num > 0) {
                int flag = constr_list(s, num, t2, nstep, &r, type);
                if (flag == 1) {
                    int *b = cm_ref(&s->cmap[num], t2);
                    if (*b & SDMD_BOND_CONNECT) *b ^= SDMD_BOND_CONNECT;
                    *b |= SDMD_CONSTRAINT_CONNECT;
                }
            } else if (t2 == 0) {
                if (strncmp(n2, "H", 1) == 0) continue;
                rd_close(&r);
                return fail(s, SDMD_EFORMAT, "residue %d: atom %s not found", AANum, n2);
            }
        }
    }
    rd_close(&r);
    bond_duplicate(s, AANum);
    constr_duplicate(s, AANum);
    return SDMD_OK;
}

/* ReadModel, Initialization.c:1076-1095 */
static int read_model(sdmd_host_system *s) {
    for (int i = 1; i <= s->naa; i++) {
        int rc = scan_aa(s, i, 0);
        if (rc) return rc;
        if (strcmp(s->aa[i].name, "GEL") == 0) continue;
        if (strcmp(s->aa[i].name, "ALA") && strcmp(s->aa[i].name, "PRO")) {
            if ((rc = scan_aa(s, i, 1))) return rc;
        }
        if (strcmp(aa_name(s, i + 1), "PRO") == 0) {
            if ((rc = scan_aa(s, i, 2))) return rc;
        }
    }
    return SDMD_OK;
}

/* one table row "T1 T2 dmin (d e)*" as parsed by Initialization.c:1492-1526 */
static int parse_table_row(sdmd_host_system *s, const char *line, int *a, int *b, double *dmin,
                           int *steps, double scale) {
    char t0[64], t1[64];
    int pos = 0;
    if (sscanf(line, "%63s%63s%n", t0, t1, &pos) < 2) return -1;
    *a = atom_model(t0);
    *b = atom_model(t1);
    if (*a < 0 || *b < 0) return fail(s, SDMD_EFORMAT, "unknown atom type in table row: %s", line);
    if (*a > *b) { int t = *a; *a = *b; *b = t; }
    const char *p = line + pos;
    char *end;
    *dmin = strtod(p, &end);
    p = end;
    *dmin *= *dmin;
    *steps = -1;
    for (;;) {
        double d = strtod(p, &end);
        if (end == p) break;
        p = end;
        double e = strtod(p, &end);
        if (end == p) break;
        p = end;
        if (*steps < 0) *steps = new_steplist(s);
        d *= d;
        steplist_push(&s->steplists[*steps], d, e * scale);
    }
    if (*steps >= 0) accumulate(&s->steplists[*steps]);
    return 0;
}

/* CreatePotentialMatrix, Initialization.c:1468-1544 */
static int create_potential_matrix(sdmd_host_system *s) {
    char path[1300], line[1024];
    snprintf(path, sizeof path, "%s/Library_Ding/InteractionPotentialTable.txt", s->datadir);
    FILE *f = fopen(path, "r");
    if (!f) return fail(s, SDMD_EIO, "cannot open %s", path);
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int a, b, steps;
        double dmin;
        int rc = parse_table_row(s, line, &a, &b, &dmin, &steps, 1.0);
        if (rc == -1) continue;
        if (rc) { fclose(f); return rc; }
        tab_t t = {dmin, 0, steps, 1};
        s->pair[a][b] = t;
        s->pair[b][a] = t;
    }
    fclose(f);
    return SDMD_OK;
}

/* ReadHB, Initialization.c:1572-1669 (table part) */
static int read_hb(sdmd_host_system *s) {
    char path[1300], line[1024], tag[256];
    snprintf(path, sizeof path, "%s/Library_Ding/HBPotentialTable.txt", s->datadir);
    FILE *f = fopen(path, "r");
    if (!f) return fail(s, SDMD_EIO, "cannot open %s", path);
    if (!fgets(line, sizeof line, f) ||
        fscanf(f, "%255s%lf%lf", tag, &s->hbpot[0], &s->hbpot[1]) != 3 ||
        fscanf(f, "%255s%lf%lf", tag, &s->hbpot[2], &s->hbpot[3]) != 3 ||
        fscanf(f, "%255s%lf%lf", tag, &s->hbpot[4], &s->hbpot[5]) != 3) {
        fclose(f);
        return fail(s, SDMD_EFORMAT, "HBPotentialTable.txt: bad header");
    }
    if (fscanf(f, "\n") < 0) { /* end of file right after the header */ }
    int cls = 0, flag = 0;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') { cls++; flag = 0; continue; }
        if (cls >= SDMD_NHBCLASS) { fclose(f); return fail(s, SDMD_EFORMAT, "too many HB classes"); }
        double scale = (cls == 1) ? s->hbpot[0] : (cls <= 5 || cls == 10) ? s->hbpot[2] : s->hbpot[4];
        int a, b, steps;
        double dmin;
        int rc = parse_table_row(s, line, &a, &b, &dmin, &steps, scale);
        if (rc == -1) continue;
        if (rc) { fclose(f); return rc; }
        tab_t t = {dmin, 0, steps, 1};
        s->hbpair[cls][a][b] = t;
        s->hbpair[cls][b][a] = t;
        if (flag == 0) { s->hbpair[cls][0][0] = t; flag = 1; }
    }
    fclose(f);
    return SDMD_OK;
}

/* HBNeighborAssign, Initialization.c:1672-1847 */
static void hb_neighbor_assign(sdmd_host_system *s, int num) {
    atomprop_t *P = s->p;
    int target = s->aa[s->pep[P[num].protein - 1].endAA + P[num].aanum].end + 1;
#define SEEK_NAME(nm) do { while (target > 1 && strcmp(P[--target].name, nm)); } while (0)
#define SEEK_TYPE(ty) do { while (target > 1 && P[--target].type != (ty)); } while (0)
    char role = 0;
    int found = 0;
    const char *n = P[num].name, *aa = P[num].aa;
    if (P[num].type == 2) { SEEK_TYPE(18); role = 'D'; found = 1; }
    else if (P[num].type == 26) { SEEK_TYPE(6); role = 'A'; found = 1; }
    else if (!strcmp(aa, "HIS")) {
        if (!strcmp(n, "HD1")) { SEEK_NAME("ND1"); role = 'D'; found = 1; }
        else if (!strcmp(n, "HE2")) { SEEK_NAME("NE2"); role = 'D'; found = 1; }
        else if (!strcmp(n, "NE2")) { SEEK_NAME("CD2"); role = 'A'; found = 1; }
    } else if (!strcmp(aa, "TYR")) {
        if (!strcmp(n, "HH")) { SEEK_NAME("OH"); role = 'D'; found = 1; }
        else if (!strcmp(n, "OH")) { SEEK_NAME("CZ"); role = 'A'; found = 1; }
    } else if (!strcmp(aa, "TRP")) {
        if (!strcmp(n, "HE1")) { SEEK_NAME("NE1"); role = 'D'; found = 1; }
    } else if (!strcmp(aa, "SER")) {
        if (!strcmp(n, "HG")) { SEEK_NAME("OG"); role = 'D'; found = 1; }
        else if (!strcmp(n, "OG")) { SEEK_NAME("CB"); role = 'A'; found = 1; }
    } else if (!strcmp(aa, "THR")) {
        if (!strcmp(n, "HG1")) { SEEK_NAME("OG1"); role = 'D'; found = 1; }
        else if (!strcmp(n, "OG1")) { SEEK_NAME("CB"); role = 'A'; found = 1; }
    } else if (!strcmp(aa, "ASN")) {
        if (!strcmp(n, "HD21") || !strcmp(n, "HD22")) { SEEK_NAME("ND2"); role = 'D'; found = 1; }
        else if (!strcmp(n, "OD1")) { SEEK_NAME("CG"); role = 'A'; found = 1; }
    } else if (!strcmp(aa, "GLN")) {
        if (!strcmp(n, "HE21") || !strcmp(n, "HE22")) { SEEK_NAME("NE2"); role = 'D'; found = 1; }
        else if (!strcmp(n, "OE1")) { SEEK_NAME("CD"); role = 'A'; found = 1; }
    } else if (!strcmp(aa, "LYS")) {
        if (!strcmp(n, "HZ1") || !strcmp(n, "HZ2") || !strcmp(n, "HZ3")) { SEEK_NAME("NZ"); role = 'D'; found = 1; }
    } else if (!strcmp(aa, "ARG")) {
        if (!strcmp(n, "HE")) { SEEK_NAME("NE"); role = 'D'; found = 1; }
        else if (!strcmp(n, "HH11") || !strcmp(n, "HH12")) { SEEK_NAME("NH1"); role = 'D'; found = 1; }
        else if (!strcmp(n, "HH21") || !strcmp(n, "HH22")) { SEEK_NAME("NH2"); role = 'D'; found = 1; }
    } else if (!strcmp(aa, "ASP")) {
        if (!strcmp(n, "OD1") || !strcmp(n, "OD2")) { SEEK_NAME("CG"); role = 'A'; found = 1; }
    } else if (!strcmp(aa, "GLU")) {
        if (!strcmp(n, "OE1") || !strcmp(n, "OE2")) { SEEK_NAME("CD"); role = 'A'; found = 1; }
    } else if (!strcmp(aa, "CYS")) {
        if (!strcmp(n, "SG")) { SEEK_NAME("CB"); role = 'S'; found = 1; }
    }
#undef SEEK_NAME
#undef SEEK_TYPE
    if (found) {
        s->hbnb[num] = target;
        s->role[num] = role;
    }
}

/* GenerateVelocity, Initialization.c:1850-1881 (no flow: constV = a = 0) */
static void generate_velocity(sdmd_host_system *s) {
    int N = s->natom;
    double *com = calloc((size_t)(N + 1) * 4, sizeof(double));
    double total = 0;
    for (int i = 1; i <= N; i++) {
        for (int n = 1; n <= 3; n++) {
            double variance = s->T / s->p[i].mass;
            double speed = mb_normal(&s->rng) * sqrt(variance) + 0;
            com[4 * i + n] = speed + 0.0 + 0.0;
            com[n] += com[4 * i + n] * s->p[i].mass;
        }
        total += s->p[i].mass;
    }
    for (int n = 1; n <= 3; n++) com[n] /= total;
    for (int i = 1; i <= N; i++) {
        if (s->nprot == 1)
            for (int n = 1; n <= 3; n++) com[4 * i + n] = com[4 * i + n] - com[n];
        s->vx[i] = com[4 * i + 1];
        s->vy[i] = com[4 * i + 2];
        s->vz[i] = com[4 * i + 3];
    }
    free(com);
}

/* ReadWall, Initialization.c:1893-1935: the smooth wall is one pseudo-atom
 * whose type and mass come from Library_Ding/Wall.txt ("Wall", then "ATOM
 * name type mass extra extra"), placed at the box centre with x = 0 for a
 * cylinder (the axis runs along x) */
static int read_wall(sdmd_host_system *s) {
    char path[1200], line[512], tag[64], name[64], rname[64], e0[64], e1[64];
    snprintf(path, sizeof path, "%s/Library_%s/Wall.txt", s->datadir, s->method);
    FILE *f = fopen(path, "r");
    if (!f) return fail(s, SDMD_EIO, "cannot open %s", path);
    double mass = 0;
    int ok = fgets(line, sizeof line, f) != NULL &&
             fscanf(f, "%63s%63s%63s%lf%63s%63s", tag, name, rname, &mass, e0, e1) == 6;
    fclose(f);
    if (!ok) return fail(s, SDMD_EFORMAT, "%s: expected a header line and an ATOM line", path);
    s->wall_type = atom_model(rname);
    if (s->wall_type <= 0) return fail(s, SDMD_EFORMAT, "%s: unknown wall atom type %s", path, rname);
    s->wall_mass = mass;
    for (int k = 1; k <= 3; k++) s->wall_pos[k] = s->box[k] / 2;
    s->wall_pos[1] = 0;                /* cylinder: the wall atom sits on the x axis */
    return SDMD_OK;
}

int sdmd_host_fail(sdmd_host_system *s, int code, const char *fmt, ...) {
    if (s && s->err && s->errlen > 0) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(s->err, (size_t)s->errlen, fmt, ap);
        va_end(ap);
    }
    return code;
}

int *sdmd_host_cm_ref(cmrow_t *row, int j) { return cm_ref(row, j); }

int sdmd_host_begin(const char *datadir, sdmd_host_system **out, char *err, int errlen) {
    if (!datadir || !out) return SDMD_EINVAL;
    sdmd_host_system *s = calloc(1, sizeof *s);
    if (!s) return SDMD_ENOMEM;
    s->err = err;
    s->errlen = errlen;
    if (err && errlen > 0) err[0] = 0;
    snprintf(s->datadir, sizeof s->datadir, "%s", datadir);
    int rc = read_parameter(s);
    if (rc) { sdmd_host_free(s); return rc; }
    *out = s;
    return SDMD_OK;
}

/* ReadModel, CreatePotentialMatrix, ReadHB and ReadWall in the order both
 * branches of InputData call them (Initialization.c:292-303, 410-423) */
int sdmd_host_finish_model(sdmd_host_system *s) {
    int rc;
    if ((rc = read_model(s))) return rc;
    if ((rc = create_potential_matrix(s))) return rc;
    if ((rc = read_hb(s))) return rc;
    if (strcmp(s->wallexist, "no") && (rc = read_wall(s))) return rc;
    return SDMD_OK;
}

static double mass_of(const sdmd_host_system *s, int lo, int hi) {
    double m = 0;
    for (int n = lo; n <= hi; n++) {
        if (strncmp(s->p[n].name, "H", 1) == 0) m++;
        else m += s->p[n].mass;
    }
    return m;
}

double sdmd_host_residue_mass(const sdmd_host_system *s, int aa) { return mass_of(s, s->aa[aa].start, s->aa[aa].end); }
double sdmd_host_chain_mass(const sdmd_host_system *s, int c) { return mass_of(s, s->pep[c].startAtom, s->pep[c].endAtom); }

double sdmd_host_charge(const sdmd_host_system *s, int i) {
    /* chargeAA[] all 0: ChargeAACheck is 0, so the template sign applies,
       and the terminal rule (!chargeAA[TER]) marks each chain's first N
       (type 18) +1 and its last O-type 26 atom -1 */
    const int c = s->p[i].protein;
    if (c >= 1 && c <= s->nprot) {
        for (int n = s->pep[c].startAtom; n <= s->pep[c].endAtom; n++)
            if (s->p[n].type == 18) { if (n == i) return 1; break; }
        for (int n = s->pep[c].endAtom; n >= s->pep[c].startAtom; n--)
            if (s->p[n].type == 26) { if (n == i) return -1; break; }
    }
    if (!strcmp(s->p[i].extra0, "+")) return 1;
    if (!strcmp(s->p[i].extra0, "-")) return -1;
    return 0;
}

int sdmd_host_load(const char *datadir, const char *gro_name, uint32_t seed,
                   sdmd_host_system **out, char *err, int errlen) {
    if (!datadir || !gro_name || !out) return SDMD_EINVAL;
    sdmd_host_system *s = NULL;
    int rc = sdmd_host_begin(datadir, &s, err, errlen);
    if (rc) return rc;
    s->seed = seed;
    sdmd_rand_seed(&s->rng, seed);
    if ((rc = read_gro(s, gro_name))) goto bad;
    if (s->natom < 2) { rc = fail(s, SDMD_EFORMAT, "no atoms"); goto bad; }
    create_cell(s);
    if ((rc = sdmd_host_finish_model(s))) goto bad;
    for (int i = 1; i <= s->natom; i++) hb_neighbor_assign(s, i);
    generate_velocity(s);
    *out = s;
    return SDMD_OK;
bad:
    sdmd_host_free(s);
    return rc;
}

/* SDEnergyMin, ToolFunctions.c:325-491 (the printing and the MiniEnergy.gro
 * snapshots are left out; the arithmetic and the rand() draws are kept) */
typedef struct { int n, cap; int *j; double *hi, *lo, *len; } inter_t;

static void inter_push(inter_t *t, int j, double hi, double lo, double len) {
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 32;
        t->j = xrealloc(t->j, sizeof(int) * t->cap);
        t->hi = xrealloc(t->hi, sizeof(double) * t->cap);
        t->lo = xrealloc(t->lo, sizeof(double) * t->cap);
        t->len = xrealloc(t->len, sizeof(double) * t->cap);
    }
    t->j[t->n] = j; t->hi[t->n] = hi; t->lo[t->n] = lo; t->len[t->n] = len;
    t->n++;
}

static int sd_energy_min(sdmd_host_system *s, long step_num) {
    int N = s->natom;
    inter_t *it = calloc((size_t)N + 1, sizeof(inter_t));
    double *c = malloc(sizeof(double) * 3 * (N + 1));
    for (int i = 1; i <= N; i++) { c[3 * i] = s->x[i]; c[3 * i + 1] = s->y[i]; c[3 * i + 2] = s->z[i]; }
    /* neighbour search on a grid of 12 A cells; pairs with d^2 <= 144 are
     * appended in increasing n, matching the reference's 1..N sweep */
    double minx = 1e300, miny = 1e300, minz = 1e300;
    for (int i = 1; i <= N; i++) {
        if (c[3 * i] < minx) minx = c[3 * i];
        if (c[3 * i + 1] < miny) miny = c[3 * i + 1];
        if (c[3 * i + 2] < minz) minz = c[3 * i + 2];
    }
    double maxx = -1e300, maxy = -1e300, maxz = -1e300;
    for (int i = 1; i <= N; i++) {
        if (c[3 * i] > maxx) maxx = c[3 * i];
        if (c[3 * i + 1] > maxy) maxy = c[3 * i + 1];
        if (c[3 * i + 2] > maxz) maxz = c[3 * i + 2];
    }
    double G = 12.5;
    if (N < 1) return SDMD_OK;
    while ((maxx - minx) / G * ((maxy - miny) / G) * ((maxz - minz) / G) > 64.0 * N + 1e6) G *= 2;
    int gx = (int)((maxx - minx) / G) + 1, gy = (int)((maxy - miny) / G) + 1, gz = (int)((maxz - minz) / G) + 1;
    long ng = (long)gx * gy * gz;
    int *head = malloc(sizeof(int) * ng), *next = malloc(sizeof(int) * (N + 1));
    for (long k = 0; k < ng; k++) head[k] = 0;
    int *gi = malloc(sizeof(int) * 3 * (N + 1));
    for (int i = N; i >= 1; i--) {   /* push front in reverse -> ascending lists */
        int a = (int)((c[3 * i] - minx) / G), b = (int)((c[3 * i + 1] - miny) / G), d = (int)((c[3 * i + 2] - minz) / G);
        gi[3 * i] = a; gi[3 * i + 1] = b; gi[3 * i + 2] = d;
        long k = ((long)d * gy + b) * gx + a;
        next[i] = head[k];
        head[k] = i;
    }
    int *cand = malloc(sizeof(int) * (N + 1));
    for (int i = 1; i <= N; i++) {
        for (int k = 0; k < s->bond[i].n; k++) {
            link_t *b = &s->bond[i].v[k];
            double hi = sqrt(b->dmax), lo = sqrt(b->dmin);
            inter_push(&it[i], b->j, hi, lo, 0.5 * (hi + lo));
        }
        for (int k = 0; k < s->constr[i].n; k++) {
            link_t *b = &s->constr[i].v[k];
            double hi = sqrt(b->dmax), lo = sqrt(b->dmin);
            inter_push(&it[i], b->j, hi, lo, 0.5 * (hi + lo));
        }
        int nc = 0;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int a = gi[3 * i] + dx, b = gi[3 * i + 1] + dy, d = gi[3 * i + 2] + dz;
                    if (a < 0 || b < 0 || d < 0 || a >= gx || b >= gy || d >= gz) continue;
                    for (int n = head[((long)d * gy + b) * gx + a]; n; n = next[n]) cand[nc++] = n;
                }
        /* restore the reference's ascending-n order */
        for (int a = 1; a < nc; a++) {
            int v = cand[a], b = a - 1;
            while (b >= 0 && cand[b] > v) { cand[b + 1] = cand[b]; b--; }
            cand[b + 1] = v;
        }
        for (int q = 0; q < nc; q++) {
            int n = cand[q];
            double d1 = c[3 * n] - c[3 * i], d2 = c[3 * n + 1] - c[3 * i + 1], d3 = c[3 * n + 2] - c[3 * i + 2];
            double l2 = d1 * d1 + d2 * d2 + d3 * d3;
            int bits = cm_get(&s->cmap[i], n);
            if (l2 <= 144 && !(bits & SDMD_BOND_CONNECT) && !(bits & SDMD_CONSTRAINT_CONNECT)) {
                int dup = 0;
                for (int k = 0; k < it[i].n; k++)
                    if (it[i].j[k] == n) { dup = 1; break; }
                if (i != n && !dup) {
                    int t1 = s->p[i].type, t2 = s->p[n].type;
                    if (t1 > t2) { int t = t1; t1 = t2; t2 = t; }
                    double lo = sqrt(s->pair[t1][t2].dmin);
                    inter_push(&it[i], n, SDMD_INFINIT, lo, lo + 0.001);
                }
            }
        }
        if (it[i].n > 510) {
            free(head); free(next); free(gi); free(cand); free(c);
            for (int k = 0; k <= N; k++) { free(it[k].j); free(it[k].hi); free(it[k].lo); free(it[k].len); }
            free(it);
            return fail(s, SDMD_ERANGE, "atom %d has %d interacting atoms (>510)", i, it[i].n);
        }
    }
    free(head); free(next); free(gi); free(cand);

    long step = 0, totalCN = 1;
    double factor = 0.1;
    while (totalCN > 0) {
        for (int i = 1; i <= N; i++) {
            double ch[3] = {0, 0, 0};
            for (int k = 0; k < it[i].n; k++) {
                int j = it[i].j[k];
                double dir[3] = {c[3 * j] - c[3 * i], c[3 * j + 1] - c[3 * i + 1], c[3 * j + 2] - c[3 * i + 2]};
                double l2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
                l2 = sqrt(l2);
                if (l2 >= it[i].hi[k] || l2 <= it[i].lo[k]) {
                    double diff = it[i].len[k] - l2;
                    double energy = rand01(&s->rng) * factor * diff;
                    ch[0] -= energy * dir[0] / it[i].len[k];
                    ch[1] -= energy * dir[1] / it[i].len[k];
                    ch[2] -= energy * dir[2] / it[i].len[k];
                }
            }
            c[3 * i] += ch[0];
            c[3 * i + 1] += ch[1];
            c[3 * i + 2] += ch[2];
        }
        totalCN = 0;
        for (int i = 1; i <= N; i++) {
            for (int k = 0; k < it[i].n; k++) {
                int j = it[i].j[k];
                if (j < i) continue;
                double dir[3] = {c[3 * j] - c[3 * i], c[3 * j + 1] - c[3 * i + 1], c[3 * j + 2] - c[3 * i + 2]};
                double l2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
                l2 = sqrt(l2);
                if (l2 >= it[i].hi[k] || l2 <= it[i].lo[k]) totalCN++;
            }
            s->x[i] = c[3 * i]; s->y[i] = c[3 * i + 1]; s->z[i] = c[3 * i + 2];
        }
        ++step;
        if (step % 1000 == 0 && factor > 0.05) factor -= 0.01;
        if (step > step_num) break;
    }
    for (int k = 0; k <= N; k++) { free(it[k].j); free(it[k].hi); free(it[k].lo); free(it[k].len); }
    free(it);
    free(c);
    return SDMD_OK;
}

/* PBC("full"), Events.c:207-228; a cylinder wall wraps x only (210-216) */
static void pbc_full(sdmd_host_system *s) {
    int dim = 3;
    if (strcmp(s->wallexist, "no") && strncmp(s->walltype, "cylinder", 1) == 0) dim = 1;
    double *c[4] = {NULL, s->x, s->y, s->z};
    for (int i = 1; i <= s->natom; i++)
        for (int n = 1; n <= dim; n++) {
            if (c[n][i] >= s->box[n]) c[n][i] -= s->box[n];
            else if (c[n][i] < 0) c[n][i] += s->box[n];
        }
}

/* LinkList("full"), Events.c:121-156 */
static int link_list_full(sdmd_host_system *s) {
    int N = s->natom;
    long ncell = (long)s->cellnum[1] * s->cellnum[2] * s->cellnum[3];
    free(s->celllist);
    s->celllist = calloc((size_t)(N + ncell + 1), sizeof(int));
    double *c[4] = {NULL, s->x, s->y, s->z};
    int *ci[4] = {NULL, s->cx, s->cy, s->cz};
    for (int i = 1; i <= N; i++) {
        for (int n = 1; n <= 3; n++) {
            ci[n][i] = (int)(c[n][i] / s->cellsize[n]);
            if (ci[n][i] == s->cellnum[n]) ci[n][i]--;
            if (ci[n][i] < 0 || ci[n][i] >= s->cellnum[n])
                return fail(s, SDMD_ERANGE, "atom %d lies outside the box after wrapping", i);
        }
        long cell = (long)ci[3][i] * s->cellnum[1] * s->cellnum[2] + (long)ci[2][i] * s->cellnum[1] + ci[1][i] + 1;
        s->celllist[i] = s->celllist[N + cell];
        s->celllist[N + cell] = i;
    }
    return SDMD_OK;
}

int sdmd_host_prepare(sdmd_host_system *s, long max_steps) {
    if (!s) return SDMD_EINVAL;
    int rc = sd_energy_min(s, max_steps);
    if (rc) return rc;
    pbc_full(s);
    return link_list_full(s);
}

int sdmd_host_normals(sdmd_host_system *s, double *out, long n, int64_t *calls_out) {
    if (!s || (!out && n > 0) || n < 0) return SDMD_EINVAL;
    for (long k = 0; k < n; k++) {
        out[k] = mb_normal(&s->rng);
        if (calls_out) calls_out[k] = s->rng.calls;
    }
    return SDMD_OK;
}

int sdmd_normal_stream(uint32_t seed, int64_t skip, double *out, long n, int64_t *calls_out) {
    if ((!out && n > 0) || n < 0 || skip < 0) return SDMD_EINVAL;
    sdmd_rand_state st;
    sdmd_rand_seed(&st, seed);
    for (int64_t k = 0; k < skip; k++) sdmd_rand_next(&st);
    for (long k = 0; k < n; k++) {
        out[k] = mb_normal(&st);
        if (calls_out) calls_out[k] = st.calls;
    }
    return SDMD_OK;
}

/* ------------------------------------------------------------------ */
/* energy: CalKinetE / CalPotenE                                       */
/* ------------------------------------------------------------------ */
static int e_type_change(int t, int direct) {
    if (direct == 1) { if (t == 1) return 3; if (t == 2) return 4; if (t == 26) return 27; }
    else { if (t == 3) return 1; if (t == 4) return 2; if (t == 27) return 26; }
    return t;
}

static const sdmd_pair *e_right_pair(const sdmd_system_desc *d, int t1, int t2) {
    const sdmd_pair *p = &d->pair[t1 * SDMD_NTYPE + t2];
    if (p->dmin > 0) return p;
    int a = e_type_change(t1, 0), b = e_type_change(t2, 0);
    if ((p = &d->pair[a * SDMD_NTYPE + t2])->dmin > 0) return p;
    if ((p = &d->pair[t1 * SDMD_NTYPE + b])->dmin > 0) return p;
    if ((p = &d->pair[a * SDMD_NTYPE + b])->dmin > 0) return p;
    return NULL;
}

static int e_hb_model(const sdmd_system_desc *d, int a1, int a2) {   /* HBModel, Models.c:212-290 */
    if (d->hb_role[a1] == 'A' || d->hb_role[a1] == 'S') { int t = a1; a1 = a2; a2 = t; }
    int o1 = e_type_change(d->type[a1], 0), o2 = e_type_change(d->type[a2], 0);
    if (o1 == 2 && o2 == 26) return 1;
    if (o1 == 2) {
        switch (d->aatype[a2]) {
        case 4: case 7: return 2; case 16: case 17: case 19: return 3; case 9: return 4;
        case 5: return 10; case 3: case 6: return -1; default: break;
        }
    } else if (o2 == 26) return 5;
    else if (o1 == 31 && o2 == 31) return 9;
    else if (o1 == 1) {
        switch (d->aatype[a2]) {
        case 4: case 7: return 6; case 16: case 17: case 19: return 7; case 9: return 8;
        case 3: case 6: return 11; default: break;
        }
    }
    return -1;
}

/* accumulated energy of FindPair (ToolFunctions.c:39-224) */
static double e_accum(const sdmd_system_desc *d, const sdmd_pair *c, double d2, double direction) {
    if (!c || c->dmin == 0) return SDMD_INFINIT;
    double tol = (direction < 0 ? 1 : -1) * SDMD_ZERO;
    if (d2 < c->dmin - SDMD_ZERO) return SDMD_INFINIT;
    for (int k = 0; k < c->nstep && d->step_d[c->step0 + k] != 0; k++)
        if (d2 < d->step_d[c->step0 + k] + tol) return d->step_acc[c->step0 + k];
    if (c->dmax > 0) return d2 < c->dmax + SDMD_ZERO ? 0 : SDMD_INFINIT;
    return SDMD_INFINIT;
}

static int e_cmap(const sdmd_system_desc *d, int i, int j) {
    int bits = 0;
    for (int k = d->cmap_off[i]; k < d->cmap_off[i + 1]; k++)
        if (d->cmap_j[k] == j) { bits = d->cmap_bits[k]; break; }
    for (int k = 0; k < SDMD_DYNSLOTS; k++)
        if (d->dyn_j[i * SDMD_DYNSLOTS + k] == j) { bits |= d->dyn_bits[i * SDMD_DYNSLOTS + k]; break; }
    return bits;
}

int sdmd_host_energy(const sdmd_system_desc *d, double *kinetic, double *potential) {
    if (!d || !kinetic || !potential) return SDMD_EINVAL;
    static const int OX[27] = {0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, -1, -1, -1, 0, 0, 1, -1, 1, -1, -1, -1, 1, 1, 1, -1};
    static const int OY[27] = {0, 0, 1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, -1, -1, -1, 1, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
    static const int OZ[27] = {0, 0, 0, 1, 1, 1, 0, 1, 0, 0, -1, -1, -1, 0, -1, 1, -1, -1, 1, 0, 0, 1, -1, -1, -1, 1, 1};
    const int N = d->natom;
    double ke = 0;
    for (int i = 1; i <= N; i++) {
        double v2 = d->vx[i] * d->vx[i] + d->vy[i] * d->vy[i] + d->vz[i] * d->vz[i];
        ke += 0.5 * d->mass[i] * v2;
    }
    double pe = 0, cut2 = d->cutoffr * d->cutoffr;
    const int *cn = d->cellnum;
    for (int i = 1; i <= N; i++) {
        int ci[3] = {d->cellx[i], d->celly[i], d->cellz[i]};
        for (int n = 0; n < 27; n++) {
            int c[3] = {ci[0] + OX[n], ci[1] + OY[n], ci[2] + OZ[n]};
            double sh[3] = {0, 0, 0};
            for (int k = 0; k < 3; k++) {
                if (c[k] < 0) { c[k] = cn[k] - 1; sh[k] = -d->box[k]; }
                else if (c[k] >= cn[k]) { c[k] = 0; sh[k] = d->box[k]; }
            }
            int cell = c[2] * cn[0] * cn[1] + c[1] * cn[0] + c[0] + 1;
            for (int j = d->celllist[N + cell]; j; j = d->celllist[j]) {
                if (j <= i) continue;
                double r[3] = {d->x[i] - (d->x[j] + sh[0]), d->y[i] - (d->y[j] + sh[1]), d->z[i] - (d->z[j] + sh[2])};
                double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                int bits = e_cmap(d, i, j);
                if (r2 > cut2 || (bits & SDMD_BOND_CONNECT)) continue;
                const sdmd_pair *p = NULL;
                int a = i, b = j;
                if (bits & SDMD_HB_CONNECT) {
                    int m = e_hb_model(d, i, j);
                    p = m >= 0 ? &d->hbpair[(m * SDMD_NTYPE) * SDMD_NTYPE] : NULL;
                    /* HBBarrier, Events.c:1097-1119 */
                    int a1 = i, a2 = j;
                    if (d->hb_role[a1] == 'A' || d->hb_role[a1] == 'S') { a1 = j; a2 = i; }
                    int o1 = e_type_change(d->type[a1], 0), o2 = e_type_change(d->type[a2], 0);
                    pe -= (o1 == 2 && o2 == 26) ? d->hb_pot[1] : (o1 != 2 && o2 != 26) ? d->hb_pot[5] : d->hb_pot[3];
                } else if (bits & SDMD_CONSTRAINT_CONNECT) {
                    for (int k = d->con_off[i]; k < d->con_off[i + 1]; k++)
                        if (d->con_j[k] == j) { p = &d->con[k]; break; }
                } else if (bits & SDMD_NEIGHBOR_CONNECT) {
                    if (d->hb_neighbor[d->hb_bond[a]] == b) { int t = a; a = b; b = t; }
                    int hbp = d->hb_bond[b];
                    int m = e_hb_model(d, b, hbp);
                    if (m >= 0)
                        p = &d->hbpair[(m * SDMD_NTYPE + e_type_change(d->type[a], 0)) * SDMD_NTYPE +
                                       e_type_change(d->type[b], 0)];
                } else {
                    p = e_right_pair(d, d->type[i], d->type[j]);
                }
                /* the direction takes the velocities of the (possibly swapped)
                 * pair against the unswapped r, ToolFunctions.c:722-736 */
                double v[3] = {d->vx[a] - d->vx[b], d->vy[a] - d->vy[b], d->vz[a] - d->vz[b]};
                double dir = r[0] * v[0] + r[1] * v[1] + r[2] * v[2];
                double acc = e_accum(d, p, r2, dir);
                if (acc < SDMD_INFINIT / 2) pe += acc;
            }
        }
    }
    *kinetic = ke;
    *potential = pe;
    return SDMD_OK;
}

/* ------------------------------------------------------------------ */
/* flattening                                                         */
/* ------------------------------------------------------------------ */
static void *flat_alloc(sdmd_host_system *s, size_t n) {
    void *p = calloc(n ? n : 1, 1);
    if (!p) abort();
    if (s->nflat < 64) s->flat_mem[s->nflat++] = p;
    return p;
}

static void flat_free(sdmd_host_system *s) {
    for (int k = 0; k < s->nflat; k++) free(s->flat_mem[k]);
    s->nflat = 0;
}

static int cmp_cm(const void *a, const void *b) {
    return ((const cm_t *)a)->j - ((const cm_t *)b)->j;
}

const sdmd_system_desc *sdmd_host_desc(sdmd_host_system *s) {
    if (!s) return NULL;
    flat_free(s);
    sdmd_system_desc *d = &s->desc;
    memset(d, 0, sizeof *d);
    int N = s->natom, N1 = N + 1;
    d->natom = N;
    for (int k = 0; k < 3; k++) {
        d->cellnum[k] = s->cellnum[k + 1];
        d->cellsize[k] = s->cellsize[k + 1];
        d->box[k] = s->box[k + 1];
    }
    d->cutoffr = s->cutoffr;
    d->temperature = s->T;
    d->thermo_freq = s->thermoF;
    d->thermostat = strcmp(s->thermostat, "no") ? 1 : 0;
    d->wall = strcmp(s->wallexist, "no") ? 1 : 0;
    d->wall_type = s->wall_type;
    d->wall_mass = s->wall_mass;
    for (int k = 0; k < 3; k++) d->wall_pos[k] = s->wall_pos[k + 1];
    d->nprotein = s->nprot;
#define ARR(T, n) ((T *)flat_alloc(s, sizeof(T) * (size_t)(n)))
    d->x = ARR(double, N1); d->y = ARR(double, N1); d->z = ARR(double, N1);
    d->vx = ARR(double, N1); d->vy = ARR(double, N1); d->vz = ARR(double, N1);
    d->mass = ARR(double, N1);
    d->type = ARR(int32_t, N1); d->aatype = ARR(int32_t, N1);
    d->protein = ARR(int32_t, N1); d->aanum = ARR(int32_t, N1);
    d->cellx = ARR(int32_t, N1); d->celly = ARR(int32_t, N1); d->cellz = ARR(int32_t, N1);
    d->hb_role = ARR(int8_t, N1); d->hb_neighbor = ARR(int32_t, N1);
    d->hb_bond = ARR(int32_t, N1); d->hb_itype = ARR(int32_t, N1);
    d->nb_partner = ARR(int32_t, N1 * SDMD_NBSLOTS);
    d->nb_partnernum = ARR(int32_t, N1); d->nb_status = ARR(int32_t, N1);
    d->dyn_j = ARR(int32_t, N1 * SDMD_DYNSLOTS); d->dyn_bits = ARR(int32_t, N1 * SDMD_DYNSLOTS);
    d->ev_partner = ARR(int32_t, N1); d->ev_counter = ARR(int64_t, N1);
    d->ev_pcounter = ARR(int64_t, N1); d->ev_time = ARR(double, N1);
    d->ev_pot = ARR(double, N1); d->ev_type = ARR(int32_t, N1); d->ev_sub = ARR(int32_t, N1);
    for (int i = 0; i <= N; i++) {
        d->x[i] = s->x[i]; d->y[i] = s->y[i]; d->z[i] = s->z[i];
        d->vx[i] = s->vx[i]; d->vy[i] = s->vy[i]; d->vz[i] = s->vz[i];
        d->mass[i] = s->p[i].mass;
        d->type[i] = s->p[i].type;
        d->aatype[i] = s->p[i].aatype;
        d->protein[i] = s->p[i].protein;
        d->aanum[i] = s->p[i].aanum;
        d->cellx[i] = s->cx[i]; d->celly[i] = s->cy[i]; d->cellz[i] = s->cz[i];
        d->hb_role[i] = (int8_t)s->role[i];
        d->hb_neighbor[i] = s->hbnb[i];
        d->ev_partner[i] = SDMD_INVALID;   /* ReadCoordinate 1019-1022 */
        d->ev_time[i] = SDMD_INFINIT;
        if (s->cont) {                     /* a continue run: the saved dynamic state */
            d->hb_bond[i] = s->hbbond[i];
            d->hb_itype[i] = s->hbitype[i];
            for (int k = 0; k < SDMD_NBSLOTS; k++) d->nb_partner[i * SDMD_NBSLOTS + k] = s->nbp[i * SDMD_NBSLOTS + k];
            d->nb_partnernum[i] = s->nbnum[i];
            d->nb_status[i] = s->nbstat[i];
            for (int k = 0; k < SDMD_DYNSLOTS; k++) {
                d->dyn_j[i * SDMD_DYNSLOTS + k] = s->dynj[i * SDMD_DYNSLOTS + k];
                d->dyn_bits[i * SDMD_DYNSLOTS + k] = s->dynb[i * SDMD_DYNSLOTS + k];
            }
            d->ev_partner[i] = s->evp[i];
            d->ev_counter[i] = s->evc[i];
            d->ev_pcounter[i] = s->evpc[i];
            d->ev_time[i] = s->evt[i];
            d->ev_pot[i] = s->evpot[i];
            d->ev_type[i] = s->evtype[i];
            d->ev_sub[i] = s->evsub[i];
        }
    }
    long ncell = (long)s->cellnum[1] * s->cellnum[2] * s->cellnum[3];
    d->ncelllist = (int32_t)(N + ncell + 1);
    d->celllist = ARR(int32_t, d->ncelllist);
    if (s->celllist) for (long k = 0; k < d->ncelllist; k++) d->celllist[k] = s->celllist[k];

    /* step pool: table step lists first, then constraint step lists */
    int total = 0;
    for (int k = 0; k < s->nsteplist; k++) total += s->steplists[k].n;
    int *base = malloc(sizeof(int) * (s->nsteplist + 1));
    d->step_d = ARR(double, total + 1); d->step_e = ARR(double, total + 1); d->step_acc = ARR(double, total + 1);
    int o = 0;
    for (int k = 0; k < s->nsteplist; k++) {
        base[k] = o;
        for (int q = 0; q < s->steplists[k].n; q++, o++) {
            d->step_d[o] = s->steplists[k].v[q].d;
            d->step_e[o] = s->steplists[k].v[q].e;
            d->step_acc[o] = s->steplists[k].v[q].acc;
        }
    }
    d->nstep_pool = total;
#define TAB(dst, src) do { (dst).dmin = (src).dmin; (dst).dmax = (src).dmax; \
        (dst).step0 = (src).steps >= 0 && (src).set ? base[(src).steps] : 0; \
        (dst).nstep = (src).steps >= 0 && (src).set ? s->steplists[(src).steps].n : 0; } while (0)
    d->pair = ARR(sdmd_pair, SDMD_NTYPE * SDMD_NTYPE);
    d->hbpair = ARR(sdmd_pair, SDMD_NHBCLASS * SDMD_NTYPE * SDMD_NTYPE);
    for (int a = 0; a < SDMD_NTYPE; a++)
        for (int b = 0; b < SDMD_NTYPE; b++) TAB(d->pair[a * SDMD_NTYPE + b], s->pair[a][b]);
    for (int c = 0; c < SDMD_NHBCLASS; c++)
        for (int a = 0; a < SDMD_NTYPE; a++)
            for (int b = 0; b < SDMD_NTYPE; b++)
                TAB(d->hbpair[(c * SDMD_NTYPE + a) * SDMD_NTYPE + b], s->hbpair[c][a][b]);
    for (int k = 0; k < 6; k++) d->hb_pot[k] = s->hbpot[k];

    /* bonds and constraints in list order */
    int nb = 0, nc = 0, ncm = 0;
    for (int i = 1; i <= N; i++) { nb += s->bond[i].n; nc += s->constr[i].n; ncm += s->cmap[i].n; }
    d->bond_off = ARR(int32_t, N + 2); d->bond_j = ARR(int32_t, nb + 1);
    d->bond_dmin = ARR(double, nb + 1); d->bond_dmax = ARR(double, nb + 1);
    d->con_off = ARR(int32_t, N + 2); d->con_j = ARR(int32_t, nc + 1); d->con = ARR(sdmd_pair, nc + 1);
    d->cmap_off = ARR(int32_t, N + 2); d->cmap_j = ARR(int32_t, ncm + 1); d->cmap_bits = ARR(int32_t, ncm + 1);
    nb = nc = ncm = 0;
    for (int i = 0; i <= N; i++) {
        d->bond_off[i] = nb; d->con_off[i] = nc; d->cmap_off[i] = ncm;
        if (i == 0) continue;
        for (int k = 0; k < s->bond[i].n; k++, nb++) {
            d->bond_j[nb] = s->bond[i].v[k].j;
            d->bond_dmin[nb] = s->bond[i].v[k].dmin;
            d->bond_dmax[nb] = s->bond[i].v[k].dmax;
        }
        for (int k = 0; k < s->constr[i].n; k++, nc++) {
            link_t *e = &s->constr[i].v[k];
            d->con_j[nc] = e->j;
            d->con[nc].dmin = e->dmin;
            d->con[nc].dmax = e->dmax;
            d->con[nc].step0 = e->steps >= 0 ? base[e->steps] : 0;
            d->con[nc].nstep = e->steps >= 0 ? s->steplists[e->steps].n : 0;
        }
        cm_t *row = malloc(sizeof(cm_t) * (s->cmap[i].n + 1));
        int m = 0;
        for (int k = 0; k < s->cmap[i].n; k++)
            if (s->cmap[i].v[k].bits) row[m++] = s->cmap[i].v[k];
        qsort(row, (size_t)m, sizeof(cm_t), cmp_cm);
        for (int k = 0; k < m; k++, ncm++) { d->cmap_j[ncm] = row[k].j; d->cmap_bits[ncm] = row[k].bits; }
        free(row);
    }
    d->bond_off[N + 1] = nb; d->con_off[N + 1] = nc; d->cmap_off[N + 1] = ncm;
    free(base);
#undef ARR
#undef TAB
    d->currenttime = s->cont ? s->currenttime : 0;
    d->seed = s->seed;
    d->rand_calls = s->rng.calls;
    return d;
}

void sdmd_host_free(sdmd_host_system *s) {
    if (!s) return;
    flat_free(s);
    if (s->bond) for (int i = 0; i <= s->natom; i++) free(s->bond[i].v);
    if (s->constr) for (int i = 0; i <= s->natom; i++) free(s->constr[i].v);
    if (s->cmap) for (int i = 0; i <= s->natom; i++) free(s->cmap[i].v);
    for (int k = 0; k < s->nsteplist; k++) free(s->steplists[k].v);
    free(s->steplists);
    free(s->bond); free(s->constr); free(s->cmap);
    free(s->p); free(s->aa); free(s->pep);
    free(s->x); free(s->y); free(s->z); free(s->vx); free(s->vy); free(s->vz);
    free(s->cx); free(s->cy); free(s->cz); free(s->role); free(s->hbnb);
    free(s->celllist);
    void *cont[] = {s->hbbond, s->hbitype, s->nbp, s->nbnum, s->nbstat, s->dynj, s->dynb, s->evp, s->evtype,
                    s->evsub, s->evc, s->evpc, s->evt, s->evpot};
    for (size_t k = 0; k < sizeof cont / sizeof cont[0]; k++) free(cont[k]);
    free(s);
}

const char *sdmd_strerror(int code) {
    switch (code) {
    case SDMD_OK: return "ok";
    case SDMD_EINVAL: return "invalid argument";
    case SDMD_EIO: return "i/o error";
    case SDMD_EFORMAT: return "bad input format";
    case SDMD_ENOMEM: return "out of memory";
    case SDMD_EUNSUPPORTED: return "unsupported option";
    case SDMD_EDEVICE: return "HIP device error";
    case SDMD_ESTATE: return "invalid state";
    case SDMD_ERANGE: return "value out of range";
    default: return "unknown error";
    }
}
