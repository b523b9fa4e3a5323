/*
 * vh_oracle.c — CPU restatement of Vahana.jl's `apply!` (test infrastructure).
 *
 * This file is the checker for the MI355X backend: tests, smoke() and the
 * CPU-baseline leg of bench.py call it; the product library never does.
 * It restates the reference algorithm with the reference's data structures
 * and per-step costs (per-target containers in insertion order, a sequential
 * slot-ordered agent loop, a full-state copy publishing the write buffer,
 * LIFO slot reuse, per-death edge removal, a sequential mapreduce fold).
 *
 * Reference paths are relative to the Vahana.jl repository root:
 *   apply! phases           src/Simulation.jl:720-821
 *   agent loop / write      src/AgentMethods.jl:159-268
 *   slots (LIFO reuse)      src/AgentMethods.jl:37-89, 430-431
 *   prepare/finish_write    src/AgentMethods.jl:271-439, src/EdgeMethods.jl:639-684
 *   containers per hint     src/EdgeMethods.jl:47-91
 *   add_edge! variants      src/EdgeMethods.jl:388-523
 *   remove_edges!           src/EdgeMethods.jl:527-599
 *   death cleanup           src/EdgeMethods.jl:606-631, 897-920
 *   accessors               src/EdgeMethods.jl:704-892, src/Edge.jl:303-335
 *   mapreduce, val4empty    src/AgentMethods.jl:533-565, src/EdgeMethods.jl:972-1001,
 *                           src/Helpers.jl:44-81
 *   AgentID codec           src/Agent.jl:30-118
 *
 * Deliberate, documented choices where the reference leaves order open:
 * non-SingleType containers are Julia Dicts keyed by AgentID; their iteration
 * order (edge mapreduce, with_edge, all_edges) is restated as ascending
 * (agent type, agent nr).
 */
#include "vh_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXA VH_MAX_AGENT_TYPES
#define MAXE VH_MAX_EDGE_TYPES

static __thread char g_err[512];
static int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
const char* or_last_error(void) { return g_err; }

/* ------------------------------------------------------------------ RNG
 * SplitMix64 finaliser (Steele, Lea, Flood 2014; constants of Vigna's
 * splitmix64.c) chained over (seed, a, b, c). */
#define GOLDEN 0x9E3779B97F4A7C15ull
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t or_hash4(uint64_t s, uint64_t a, uint64_t b, uint64_t c) {
    uint64_t h = mix64(s + GOLDEN);
    h = mix64(h + a + GOLDEN);
    h = mix64(h + b + GOLDEN);
    return mix64(h + c + GOLDEN);
}
static double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }
static float u01f(uint64_t h) { return (float)(h >> 40) * 0x1.0p-24f; }

/* AgentID codec, src/Agent.jl:63-115 (rank 0) */
static uint64_t agent_id(uint64_t type, uint64_t nr) { return (type << VH_SHIFT_TYPE) | nr; }
static int type_nr(uint64_t id) { return (int)(id >> VH_SHIFT_TYPE); }
static uint64_t agent_nr(uint64_t id) { return id & VH_NR_MASK; }

/* ------------------------------------------------------------------ model */
typedef struct { uint32_t bytes, hints; } atype_t;
typedef struct { uint32_t bytes, hints; int32_t target; int64_t size; } etype_t;
struct or_model { int na, ne; atype_t a[MAXA + 1]; etype_t e[MAXE]; };

int or_model_create(or_model** out) {
    *out = (or_model*)calloc(1, sizeof(or_model));
    return *out ? VH_OK : fail(VH_ERR_OOM, "oom");
}
int or_model_destroy(or_model* m) { free(m); return VH_OK; }

/* register_agenttype! src/ModelTypes.jl:81-108 */
int or_register_agenttype(or_model* m, uint32_t hints, uint32_t bytes, uint8_t* id) {
    if (m->na + 1 > MAXA) return fail(VH_ERR_ASSERT, "maximal number of types already registered");
    if (hints & ~(VH_HINT_IMMORTAL | VH_HINT_INDEPENDENT))
        return fail(VH_ERR_ASSERT, "unknown agent type hint");
    m->na++;
    m->a[m->na].bytes = bytes;
    m->a[m->na].hints = hints;
    *id = (uint8_t)m->na;
    return VH_OK;
}

/* register_edgetype! src/ModelTypes.jl:158-230 */
int or_register_edgetype(or_model* m, uint32_t hints, uint32_t bytes, int32_t target,
                         int64_t size, uint8_t* id) {
    if (m->ne >= MAXE) return fail(VH_ERR_ASSERT, "too many edge types");
    if (hints & VH_HINT_NUMEDGESONLY) hints |= VH_HINT_STATELESS | VH_HINT_IGNOREFROM;
    if (hints & VH_HINT_HASEDGEONLY)
        hints |= VH_HINT_STATELESS | VH_HINT_IGNOREFROM | VH_HINT_SINGLEEDGE;
    if ((hints & VH_HINT_SINGLETYPE) && target <= 0)
        return fail(VH_ERR_ASSERT, ":SingleType needs the target keyword");
    if (target > 0) hints |= VH_HINT_SINGLETYPE;
    if (target > m->na) return fail(VH_ERR_UNREGISTERED, "unregistered target type");
    if ((hints & VH_HINT_SINGLETYPE) && (hints & VH_HINT_SINGLEEDGE) &&
        !((hints & VH_HINT_STATELESS) && (hints & VH_HINT_IGNOREFROM)))
        return fail(VH_ERR_ASSERT,
                    "The hints :SingleEdge and :SingleType can be only combined when the "
                    "type has also the hints :Stateless and :IgnoreFrom.");
    etype_t* e = &m->e[m->ne];
    e->bytes = (hints & VH_HINT_STATELESS) ? 0 : bytes;
    e->hints = hints;
    e->target = target > 0 ? target : 0;
    e->size = size;
    *id = (uint8_t)m->ne++;
    return VH_OK;
}
int or_edgetype_hints(or_model* m, uint8_t et, uint32_t* hints) {
    if (et >= m->ne) return fail(VH_ERR_UNREGISTERED, "unregistered edge type");
    *hints = m->e[et].hints;
    return VH_OK;
}

/* ------------------------------------------------------------ containers */
typedef struct { uint64_t* v; uint64_t n, cap; } u64vec;
static void u64_push(u64vec* a, uint64_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 4;
        a->v = (uint64_t*)realloc(a->v, a->cap * sizeof(uint64_t));
    }
    a->v[a->n++] = x;
}

/* AgentReadWrite src/Simulation.jl:13-35 */
typedef struct {
    uint8_t* state; uint64_t len;      /* state vector; len = length(state) */
    uint8_t* died;                     /* mortal only, len entries          */
    u64vec reuse;                      /* reuseable slot numbers            */
    uint64_t cap;                      /* allocated slots (host amortisation) */
} rw_t;
typedef struct { rw_t read, write; uint64_t nextid; int prepared, writeable; uint64_t last_change; } agents_t;

/* One per-target container (the element of the Vector/Dict field).  For
 * Stateless+IgnoreFrom layouts `n` is the Int64 count (or the Bool). */
typedef struct { uint32_t n, cap; uint64_t* from; uint8_t* st; uint8_t present; } cont_t;
typedef struct { cont_t* rows[MAXA + 1]; uint64_t nrows[MAXA + 1]; } field_t;

typedef struct {
    field_t* read; field_t* write;
    int readable, writeable, add_existing;
    /* agentsontarget: source id -> targets (src/EdgeMethods.jl:375-384) */
    u64vec* aot[MAXA + 1]; uint8_t* aot_has[MAXA + 1]; uint64_t aot_n[MAXA + 1];
    uint64_t last_change;
} edges_t;

struct or_sim {
    or_model* m;                       /* -> mown */
    or_model mown;                     /* own copy: the model may be destroyed first */
    agents_t a[MAXA + 1];
    edges_t e[MAXE];
    int initialized, intransition;
    uint64_t num_transitions;
    const vh_apply_desc* d;            /* desc of the running apply! */
    uint8_t params[256];
    /* param / get_global values: raw 8 bytes, vh_dtype, set mask per kind */
    uint64_t vals[2][VH_MAX_VALUES];
    uint8_t vdt[2][VH_MAX_VALUES];
    uint32_t vset[2];
};

#define ETH(s, et) ((s)->m->e[et].hints)
#define HAS(h, x) (((h) & (x)) != 0)
static int is_count(uint32_t h) { return HAS(h, VH_HINT_STATELESS) && HAS(h, VH_HINT_IGNOREFROM); }
static int mortal_t(or_sim* s, int t) { return !HAS(s->m->a[t].hints, VH_HINT_IMMORTAL); }

static void cont_free(cont_t* c) { free(c->from); free(c->st); c->from = NULL; c->st = NULL; c->n = c->cap = 0; }
static field_t* field_new(void) { return (field_t*)calloc(1, sizeof(field_t)); }
static void field_free(field_t* f) {
    if (!f) return;
    for (int t = 0; t <= MAXA; t++) {
        for (uint64_t i = 0; i < f->nrows[t]; i++) cont_free(&f->rows[t][i]);
        free(f->rows[t]);
    }
    free(f);
}
static field_t* field_copy(const field_t* f, uint32_t bytes) {
    field_t* g = field_new();
    for (int t = 0; t <= MAXA; t++) {
        g->nrows[t] = f->nrows[t];
        if (!f->nrows[t]) continue;
        g->rows[t] = (cont_t*)calloc(f->nrows[t], sizeof(cont_t));
        for (uint64_t i = 0; i < f->nrows[t]; i++) {
            const cont_t* a = &f->rows[t][i];
            cont_t* b = &g->rows[t][i];
            b->n = a->n; b->present = a->present; b->cap = a->n;
            if (a->from && a->n) {
                b->from = (uint64_t*)malloc(a->n * 8);
                memcpy(b->from, a->from, a->n * 8);
            }
            if (a->st && a->n && bytes) {
                b->st = (uint8_t*)malloc((size_t)a->n * bytes);
                memcpy(b->st, a->st, (size_t)a->n * bytes);
            }
        }
    }
    return g;
}
static cont_t* field_get(field_t* f, uint64_t to) {
    int t = type_nr(to);
    uint64_t nr = agent_nr(to);
    if (t < 1 || t > MAXA || nr == 0 || nr > f->nrows[t]) return NULL;
    cont_t* c = &f->rows[t][nr - 1];
    return c->present ? c : NULL;
}
static cont_t* field_get_mut(field_t* f, uint64_t to) {
    int t = type_nr(to);
    uint64_t nr = agent_nr(to);
    if (nr > f->nrows[t]) {
        uint64_t nn = f->nrows[t] ? f->nrows[t] : 16;
        while (nn < nr) nn *= 2;
        f->rows[t] = (cont_t*)realloc(f->rows[t], nn * sizeof(cont_t));
        memset(f->rows[t] + f->nrows[t], 0, (nn - f->nrows[t]) * sizeof(cont_t));
        f->nrows[t] = nn;
    }
    cont_t* c = &f->rows[t][nr - 1];
    c->present = 1;
    return c;
}
static void cont_push(cont_t* c, uint64_t from, const uint8_t* st, uint32_t bytes, int keep_from) {
    if (c->n == c->cap) {
        c->cap = c->cap ? c->cap * 2 : 2;
        if (keep_from) c->from = (uint64_t*)realloc(c->from, c->cap * 8);
        if (bytes) c->st = (uint8_t*)realloc(c->st, (size_t)c->cap * bytes);
    }
    if (keep_from) c->from[c->n] = from;
    if (bytes) memcpy(c->st + (size_t)c->n * bytes, st, bytes);
    c->n++;
}

/* ------------------------------------------------------------ simulation */
int or_sim_create(or_model* m, or_sim** out) {
    or_sim* s = (or_sim*)calloc(1, sizeof(or_sim));
    if (!s) return fail(VH_ERR_OOM, "oom");
    s->mown = *m;
    s->m = &s->mown;
    for (int t = 1; t <= m->na; t++) s->a[t].nextid = 1;
    for (int e = 0; e < m->ne; e++) {
        s->e[e].read = field_new();
        s->e[e].write = s->e[e].read;      /* alias until a prepare_write! */
    }
    *out = s;
    return VH_OK;
}
static void aot_clear(edges_t* x) {
    for (int t = 0; t <= MAXA; t++) {
        for (uint64_t i = 0; i < x->aot_n[t]; i++) free(x->aot[t][i].v);
        free(x->aot[t]); free(x->aot_has[t]);
        x->aot[t] = NULL; x->aot_has[t] = NULL; x->aot_n[t] = 0;
    }
}
static void rw_free(rw_t* r) { free(r->state); free(r->died); free(r->reuse.v); memset(r, 0, sizeof *r); }
int or_sim_destroy(or_sim* s) {
    if (!s) return VH_OK;
    for (int t = 1; t <= s->m->na; t++) { rw_free(&s->a[t].read); rw_free(&s->a[t].write); }
    for (int e = 0; e < s->m->ne; e++) {
        if (s->e[e].write != s->e[e].read) field_free(s->e[e].write);
        field_free(s->e[e].read);
        aot_clear(&s->e[e]);
    }
    free(s);
    return VH_OK;
}
static void rw_copy(rw_t* dst, const rw_t* src, uint32_t bytes, int mortal) {
    dst->len = src->len;
    dst->cap = src->len;
    dst->state = (uint8_t*)malloc(src->len * bytes + 1);
    if (src->len && bytes) memcpy(dst->state, src->state, src->len * bytes);
    dst->died = NULL;
    if (mortal) {
        dst->died = (uint8_t*)malloc(src->len + 1);
        if (src->len) memcpy(dst->died, src->died, src->len);
    }
    dst->reuse.n = src->reuse.n; dst->reuse.cap = src->reuse.n;
    dst->reuse.v = (uint64_t*)malloc(src->reuse.n * 8 + 8);
    if (src->reuse.n) memcpy(dst->reuse.v, src->reuse.v, src->reuse.n * 8);
}
/* copy_simulation src/Simulation.jl:500-510 (deepcopy) */
int or_sim_copy(or_sim* s, or_sim** out) {
    or_sim* c = (or_sim*)calloc(1, sizeof(or_sim));
    *c = *s;
    c->m = &c->mown;
    for (int t = 1; t <= s->m->na; t++) {
        rw_copy(&c->a[t].read, &s->a[t].read, s->m->a[t].bytes, mortal_t(s, t));
        rw_copy(&c->a[t].write, &s->a[t].write, s->m->a[t].bytes, mortal_t(s, t));
    }
    for (int e = 0; e < s->m->ne; e++) {
        c->e[e].read = field_copy(s->e[e].read, s->m->e[e].bytes);
        c->e[e].write = s->e[e].write == s->e[e].read ? c->e[e].read
                                                      : field_copy(s->e[e].write, s->m->e[e].bytes);
        for (int t = 0; t <= MAXA; t++) {
            c->e[e].aot_n[t] = s->e[e].aot_n[t];
            if (!s->e[e].aot_n[t]) { c->e[e].aot[t] = NULL; c->e[e].aot_has[t] = NULL; continue; }
            c->e[e].aot[t] = (u64vec*)calloc(s->e[e].aot_n[t], sizeof(u64vec));
            c->e[e].aot_has[t] = (uint8_t*)malloc(s->e[e].aot_n[t]);
            memcpy(c->e[e].aot_has[t], s->e[e].aot_has[t], s->e[e].aot_n[t]);
            for (uint64_t i = 0; i < s->e[e].aot_n[t]; i++)
                for (uint64_t j = 0; j < s->e[e].aot[t][i].n; j++)
                    u64_push(&c->e[e].aot[t][i], s->e[e].aot[t][i].v[j]);
        }
    }
    *out = c;
    return VH_OK;
}

/* ------------------------------------------------------------ agents */
static void rw_resize(rw_t* w, uint64_t len, uint32_t bytes, int mortal) {
    if (len > w->cap || (mortal && !w->died)) {
        uint64_t nc = w->cap > len ? w->cap : (w->cap * 2 > len ? w->cap * 2 : len);
        w->state = (uint8_t*)realloc(w->state, nc * bytes + 1);
        if (mortal) w->died = (uint8_t*)realloc(w->died, nc + 1);
        w->cap = nc;
    }
    if (len > w->len) {
        memset(w->state + w->len * bytes, 0, (len - w->len) * bytes);
        if (mortal) memset(w->died + w->len, 0, len - w->len);
    }
    w->len = len;
}

/* _get_next_id + add_agent! src/AgentMethods.jl:37-89 */
static uint64_t add_agent_impl(or_sim* s, int t, const uint8_t* st) {
    agents_t* A = &s->a[t];
    uint32_t bytes = s->m->a[t].bytes;
    int mortal = mortal_t(s, t);
    int indep = HAS(s->m->a[t].hints, VH_HINT_INDEPENDENT);
    uint64_t nr;
    if (mortal && A->read.reuse.n > 0) {
        nr = A->read.reuse.v[--A->read.reuse.n];            /* pop! (LIFO) */
        A->write.died[nr - 1] = 0;
    } else {
        nr = A->nextid++;
        if (nr > A->write.len) rw_resize(&A->write, nr, bytes, mortal);   /* resize! */
        if (mortal) A->write.died[nr - 1] = 0;
    }
    if (indep && nr <= A->read.len) memcpy(A->read.state + (nr - 1) * bytes, st, bytes);
    else memcpy(A->write.state + (nr - 1) * bytes, st, bytes);
    return agent_id((uint64_t)t, nr);
}

int or_add_agents(or_sim* s, uint8_t t, const void* states, uint64_t n, uint64_t* ids) {
    if (t < 1 || t > s->m->na) return fail(VH_ERR_UNREGISTERED, "unregistered agent type");
    if (s->initialized && !s->intransition)
        return fail(VH_ERR_ASSERT, "You can call add_agent! only in the initialization phase "
                                   "(until `finish_init!` is called) or within a transition function");
    uint32_t bytes = s->m->a[t].bytes;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t id = add_agent_impl(s, t, (const uint8_t*)states + i * bytes);
        if (ids) ids[i] = id;
    }
    return VH_OK;
}

/* ------------------------------------------------------------ edges */
static void aot_push(edges_t* x, uint64_t from, uint64_t to) {
    int t = type_nr(from);
    uint64_t nr = agent_nr(from);
    if (nr > x->aot_n[t]) {
        uint64_t nn = x->aot_n[t] ? x->aot_n[t] : 16;
        while (nn < nr) nn *= 2;
        x->aot[t] = (u64vec*)realloc(x->aot[t], nn * sizeof(u64vec));
        memset(x->aot[t] + x->aot_n[t], 0, (nn - x->aot_n[t]) * sizeof(u64vec));
        x->aot_has[t] = (uint8_t*)realloc(x->aot_has[t], nn);
        memset(x->aot_has[t] + x->aot_n[t], 0, nn - x->aot_n[t]);
        x->aot_n[t] = nn;
    }
    x->aot_has[t][nr - 1] = 1;
    u64_push(&x->aot[t][nr - 1], to);
}

static int valid_target(or_sim* s, int et, uint64_t to) {
    int t = type_nr(to);
    if (t < 1 || t > s->m->na || agent_nr(to) == 0) return 0;
    if (HAS(ETH(s, et), VH_HINT_SINGLETYPE) && t != s->m->e[et].target) return 0;
    return 1;
}

/* add_edge! src/EdgeMethods.jl:388-523 with _can_add :257-294 and
 * _push_agentsontarget :375-384 */
static int add_edge_impl(or_sim* s, int et, uint64_t from, uint64_t to, const uint8_t* st) {
    edges_t* X = &s->e[et];
    uint32_t h = ETH(s, et), bytes = s->m->e[et].bytes;
    if (s->initialized && !s->intransition)
        return fail(VH_ERR_ASSERT, "You can call add_edge! only in the initialization phase "
                                   "(until `finish_init!` is called) or within a transition function");
    if (s->initialized && !X->writeable)
        return fail(VH_ERR_ASSERT, "edge type %d must be in the `write` argument of the "
                                   "transition function.", et);
    if (!valid_target(s, et, to))
        return fail(VH_ERR_ASSERT, "The :SingleType hint is set / the target id is invalid");
    int ignorefrom = HAS(h, VH_HINT_IGNOREFROM), stateless = HAS(h, VH_HINT_STATELESS);
    int singleedge = HAS(h, VH_HINT_SINGLEEDGE), singletype = HAS(h, VH_HINT_SINGLETYPE);
    if (!ignorefrom && (type_nr(from) < 1 || type_nr(from) > s->m->na || agent_nr(from) == 0))
        return fail(VH_ERR_ASSERT, "invalid source id");
    if (singleedge && !singletype && !(ignorefrom && stateless)) {
        cont_t* c = field_get(X->write, to);
        if (c && c->n > 0) {
            int same = 1;
            if (!ignorefrom && c->from[0] != from) same = 0;
            if (!stateless && memcmp(c->st, st, bytes) != 0) same = 0;
            if (!same)
                return fail(VH_ERR_ASSERT, "An edge has already been added to the agent with the "
                                           "id %llu (and the edgetype hints containing the "
                                           ":SingleEdge hint).", (unsigned long long)to);
        }
    }
    if (!ignorefrom) {
        int all_immortal = 1;
        for (int t = 1; t <= s->m->na; t++) if (mortal_t(s, t)) all_immortal = 0;
        if (!all_immortal && mortal_t(s, type_nr(from))) aot_push(X, from, to);
    }
    cont_t* c = field_get_mut(X->write, to);
    if (is_count(h)) {
        if (singleedge) c->n = 1; else c->n += 1;
    } else if (singleedge) {
        c->n = 0;
        cont_push(c, from, st, bytes, !ignorefrom);
    } else {
        cont_push(c, from, st, bytes, !ignorefrom);
    }
    return VH_OK;
}

int or_add_edges(or_sim* s, uint8_t et, const uint64_t* from, const uint64_t* to,
                 const void* states, uint64_t n) {
    if (et >= s->m->ne) return fail(VH_ERR_UNREGISTERED, "unregistered edge type");
    uint32_t bytes = s->m->e[et].bytes;
    static const uint8_t zero[64];
    for (uint64_t i = 0; i < n; i++) {
        const uint8_t* st = (states && bytes) ? (const uint8_t*)states + i * bytes : zero;
        int r = add_edge_impl(s, et, from ? from[i] : 0, to[i], st);
        if (r) return r;
    }
    return VH_OK;
}

/* remove_edges!(sim, to, T) src/EdgeMethods.jl:527-549 */
static void clear_target(or_sim* s, int et, uint64_t to) {
    field_t* f = s->e[et].write;
    if (HAS(ETH(s, et), VH_HINT_SINGLETYPE)) {
        cont_t* c = field_get_mut(f, to);    /* field[nr] = zero(CT): assigned, empty */
        c->n = 0;
    } else {
        cont_t* c = field_get(f, to);        /* delete!(field, to) */
        if (c) { c->n = 0; c->present = 0; }
    }
}

/* remove_edges!(sim, from, to, T) src/EdgeMethods.jl:552-599 */
static void remove_from_to(or_sim* s, int et, uint64_t from, uint64_t to) {
    uint32_t h = ETH(s, et), bytes = s->m->e[et].bytes;
    field_t* f = s->e[et].write;
    int t = type_nr(to);
    uint64_t nr = agent_nr(to);
    int singletype = HAS(h, VH_HINT_SINGLETYPE);
    cont_t* c = NULL;
    if (singletype) { if (t >= 1 && t <= MAXA && nr >= 1 && nr <= f->nrows[t]) c = &f->rows[t][nr - 1]; }
    else c = field_get(f, to);
    if (!c) return;
    if (HAS(h, VH_HINT_SINGLEEDGE)) {
        if (c->n > 0 && c->from[0] == from) clear_target(s, et, to);
    } else {
        if (!c->present || c->n == 0) return;
        uint32_t k = 0;
        for (uint32_t i = 0; i < c->n; i++) {
            if (c->from[i] == from) continue;
            c->from[k] = c->from[i];
            if (bytes) memmove(c->st + (size_t)k * bytes, c->st + (size_t)i * bytes, bytes);
            k++;
        }
        c->n = k;
    }
}

/* _remove_edges_agent_target! src/EdgeMethods.jl:897-920 */
static void remove_agent_target(or_sim* s, int et, uint64_t id) {
    uint32_t h = ETH(s, et);
    field_t* f = s->e[et].write;
    if (HAS(h, VH_HINT_SINGLETYPE)) {
        if (type_nr(id) != s->m->e[et].target) return;
        uint64_t nr = agent_nr(id);
        if (f->nrows[type_nr(id)] >= nr) { cont_t* c = &f->rows[type_nr(id)][nr - 1]; c->n = 0; c->present = 1; }
    } else {
        cont_t* c = field_get(f, id);
        if (c) { c->n = 0; c->present = 0; }
    }
}

/* _remove_edges_agent_source! src/EdgeMethods.jl:606-631 */
static void remove_agent_source(or_sim* s, int et, const u64vec* died) {
    edges_t* X = &s->e[et];
    if (HAS(ETH(s, et), VH_HINT_IGNOREFROM)) return;
    for (uint64_t i = 0; i < died->n; i++) {
        uint64_t f = died->v[i];
        int t = type_nr(f);
        uint64_t nr = agent_nr(f);
        if (nr == 0 || nr > X->aot_n[t] || !X->aot_has[t][nr - 1]) continue;
        u64vec* l = &X->aot[t][nr - 1];
        for (uint64_t j = 0; j < l->n; j++) remove_from_to(s, et, f, l->v[j]);
        free(l->v); memset(l, 0, sizeof *l);
        X->aot_has[t][nr - 1] = 0;
    }
}

/* ------------------------------------------------------------ finish_write */
/* edges (or counts) held by the containers of a field */
static uint64_t field_edges(const field_t* f) {
    uint64_t c = 0;
    if (!f) return 0;
    for (int t = 0; t <= MAXA; t++)
        for (uint64_t r = 0; f->rows[t] && r < f->nrows[t]; r++) c += f->rows[t][r].n;
    return c;
}

/* agent finish_write! src/AgentMethods.jl:300-439 */
static void finish_write_agents(or_sim* s, int t) {
    agents_t* A = &s->a[t];
    uint32_t bytes = s->m->a[t].bytes;
    int mortal = mortal_t(s, t), indep = HAS(s->m->a[t].hints, VH_HINT_INDEPENDENT);
    int must_copy = A->write.len > A->read.len || !s->initialized || !indep;
    if (mortal) {
        u64vec aids = {0, 0, 0};
        for (uint64_t i = 0; i < A->write.reuse.n; i++)
            u64_push(&aids, agent_id((uint64_t)t, A->write.reuse.v[i]));
        if (aids.n) {
            /* network_changed: an edge type that lost edges to the deaths
             * gets last_change = num_transitions (src/AgentMethods.jl:347-353;
             * there the flag is set when the dead agent had a container or
             * a source entry, here when edges were removed) */
            uint64_t before[MAXE];
            for (int et = 0; et < s->m->ne; et++) before[et] = field_edges(s->e[et].write);
            for (uint64_t i = 0; i < aids.n; i++)
                for (int et = 0; et < s->m->ne; et++) remove_agent_target(s, et, aids.v[i]);
            for (int et = 0; et < s->m->ne; et++) remove_agent_source(s, et, &aids);
            for (int et = 0; et < s->m->ne; et++)
                if (field_edges(s->e[et].write) != before[et]) s->e[et].last_change = s->num_transitions;
        }
        free(aids.v);
    }
    if (bytes && must_copy) {
        if (indep && s->initialized && A->read.len)
            memcpy(A->write.state, A->read.state, A->read.len * bytes);
        free(A->read.state);                       /* new window + memcpy */
        A->read.state = (uint8_t*)malloc(A->write.len * bytes + 1);
        memcpy(A->read.state, A->write.state, A->write.len * bytes);
        A->read.len = A->write.len;
        A->read.cap = A->write.len;
    } else if (!bytes) {
        A->read.len = A->write.len;
    }
    if (mortal) {
        free(A->read.died);
        A->read.died = (uint8_t*)malloc(A->write.len + 1);
        memcpy(A->read.died, A->write.died, A->write.len);
    }
    for (uint64_t i = 0; i < A->write.reuse.n; i++) u64_push(&A->read.reuse, A->write.reuse.v[i]);
    A->write.reuse.n = 0;
    A->last_change = s->num_transitions;
    A->writeable = 0;
}

/* edge finish_write! src/EdgeMethods.jl:680-684 */
static void finish_write_edges(or_sim* s, int et) {
    edges_t* X = &s->e[et];
    if (X->write != X->read) { field_free(X->read); X->read = X->write; }
    X->last_change = s->num_transitions;
    X->writeable = 0;
}

int or_last_change(or_sim* s, int kind, uint8_t t, uint64_t* out) {
    if (kind == 0) {
        if (t < 1 || t > s->m->na) return fail(VH_ERR_UNREGISTERED, "unregistered agent type");
        *out = s->a[t].last_change;
    } else {
        if (t >= s->m->ne) return fail(VH_ERR_UNREGISTERED, "unregistered edge type");
        *out = s->e[t].last_change;
    }
    return VH_OK;
}
int or_set_last_change(or_sim* s, int kind, uint8_t t, uint64_t v) {
    if (kind == 0) {
        if (t < 1 || t > s->m->na) return fail(VH_ERR_UNREGISTERED, "unregistered agent type");
        s->a[t].last_change = v;
    } else {
        if (t >= s->m->ne) return fail(VH_ERR_UNREGISTERED, "unregistered edge type");
        s->e[t].last_change = v;
    }
    return VH_OK;
}

/* finish_init! src/Simulation.jl:403-476 (one process: no distribute!) */
int or_finish_init(or_sim* s) {
    if (s->initialized) return fail(VH_ERR_ASSERT, "finish_init! was already called");
    for (int t = 1; t <= s->m->na; t++) finish_write_agents(s, t);
    for (int et = 0; et < s->m->ne; et++) finish_write_edges(s, et);
    s->initialized = 1;
    s->num_transitions = 1;
    return VH_OK;
}

/* ------------------------------------------------------- body accessors */
typedef struct {
    or_sim* s; const vh_apply_desc* d; int err;
    uint64_t step; int call;
} ctx_t;

static cont_t* rd_cont(ctx_t* c, int et, uint64_t to) {
    if (et < 0 || et >= c->s->m->ne) { c->err = fail(VH_ERR_ARG, "unbound edge role"); return NULL; }
    if (!c->s->e[et].readable) {
        c->err = fail(VH_ERR_ASSERT, "You try to access an edge of type %d but the type is not in "
                                     "the `read` argument of the apply! function.", et);
        return NULL;
    }
    return field_get(c->s->e[et].read, to);
}
static uint32_t num_edges_of(ctx_t* c, int et, uint64_t to) {
    cont_t* k = rd_cont(c, et, to);
    return k ? k->n : 0;
}
/* agentstate / agentstate_flexible src/AgentMethods.jl:91-123 */
static const uint8_t* agentstate_of(ctx_t* c, uint64_t id) {
    int t = type_nr(id);
    if (t < 1 || t > c->s->m->na) { c->err = fail(VH_ERR_ASSERT, "invalid agent id"); return NULL; }
    agents_t* A = &c->s->a[t];
    if (!A->prepared) {
        c->err = fail(VH_ERR_ASSERT, "agent type %d must be in the `read` argument of the "
                                     "transition function.", t);
        return NULL;
    }
    uint64_t nr = agent_nr(id);
    if (nr == 0 || nr > A->read.len) { c->err = fail(VH_ERR_ASSERT, "agent nr out of range"); return NULL; }
    if (mortal_t(c->s, t) && A->read.died[nr - 1]) {
        c->err = fail(VH_ERR_ASSERT, "agentstate was requested for agent %llu that has been removed",
                      (unsigned long long)id);
        return NULL;
    }
    static const uint8_t zero[64];
    return c->s->m->a[t].bytes ? A->read.state + (nr - 1) * c->s->m->a[t].bytes : zero;
}
static void ctx_add_edge(ctx_t* c, int et, uint64_t from, uint64_t to, const uint8_t* st) {
    if (et < 0 || et >= c->s->m->ne) { c->err = fail(VH_ERR_ARG, "unbound edge role"); return; }
    int r = add_edge_impl(c->s, et, from, to, st);
    if (r) c->err = r;
}
static uint64_t ctx_add_agent(ctx_t* c, int t, const uint8_t* st) {
    if (t < 1 || t > c->s->m->na) { c->err = fail(VH_ERR_ARG, "unbound agent role"); return 0; }
    if (!c->s->a[t].writeable) {
        c->err = fail(VH_ERR_ASSERT, "agent type %d must be in the `write` argument", t);
        return 0;
    }
    return add_agent_impl(c->s, t, st);
}

/* remove_edges! inside a transition (src/EdgeMethods.jl:527-599) with
 * _can_remove_edges (:101-121); `all`: remove_edges!(sim, to, T) */
static void ctx_remove_edges(ctx_t* c, int et, uint64_t from, uint64_t to, int all) {
    if (et < 0 || et >= c->s->m->ne) { c->err = fail(VH_ERR_ARG, "unbound edge role"); return; }
    edges_t* X = &c->s->e[et];
    if (!X->writeable) {
        c->err = fail(VH_ERR_ASSERT, "Edge of type %d can not removed, as the type is not in the "
                                     "`write` argument of the transition function.", et);
        return;
    }
    if (!X->add_existing) {
        c->err = fail(VH_ERR_ASSERT, "Edge of type %d can not removed, as the type is not in the "
                                     "`add_existing` argument of the transition function.", et);
        return;
    }
    if (!valid_target(c->s, et, to)) { c->err = fail(VH_ERR_ASSERT, "invalid edge target"); return; }
    if (all) { clear_target(c->s, et, to); return; }
    if (HAS(ETH(c->s, et), VH_HINT_IGNOREFROM)) {
        c->err = fail(VH_ERR_ASSERT, "remove_edges! with a source agent is not defined for "
                                     "edgetypes with the :IgnoreFrom hint.");
        return;
    }
    remove_from_to(c->s, et, from, to);
}

/* ------------------------------------------------------------ model bodies
 * Each returns 1 (state in `out`) or 0 (`nothing`: the agent dies). */
typedef int (*body_fn)(ctx_t*, const uint8_t* st, uint64_t id, uint8_t* out);

static int64_t rd_i64(const uint8_t* p) { int64_t v; memcpy(&v, p, 8); return v; }
static void wr_i64(uint8_t* p, int64_t v) { memcpy(p, &v, 8); }

static int body_noop(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)c; (void)id; (void)st; (void)out;
    return 1;
}
/* create_sum_state_neighbors test/core.jl:71-83 */
static int body_sum_neighbors(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)st;
    int et = c->d->role_edge[0];
    cont_t* k = rd_cont(c, et, id);
    int64_t sum = 0;
    if (k) {
        if (HAS(ETH(c->s, et), VH_HINT_IGNOREFROM)) { c->err = fail(VH_ERR_ASSERT, "neighborids is not defined for the hint combination"); return 1; }
        for (uint32_t i = 0; i < k->n; i++) {
            const uint8_t* ns = agentstate_of(c, k->from[i]);
            if (!ns) return 1;
            sum += rd_i64(ns);
        }
    }
    wr_i64(out, sum);
    return 1;
}
static int pred(ctx_t* c, const uint8_t* st, uint64_t id) {
    const vh_test_params* p = (const vh_test_params*)c->s->params;
    int64_t foo = c->s->m->a[type_nr(id)].bytes >= 8 ? rd_i64(st) : 0;
    switch (p->cmp) {
    case VH_CMP_ALWAYS: return 1;
    case VH_CMP_NEVER: return 0;
    case VH_CMP_LT: return foo < p->value;
    case VH_CMP_GE: return foo >= p->value;
    case VH_CMP_EQ: return foo == p->value;
    case VH_CMP_ODD: return (foo % 2) != 0;
    case VH_CMP_EVEN: return (foo % 2) == 0;
    case VH_CMP_NUM_EDGES_EQ: return (int64_t)num_edges_of(c, c->d->role_edge[0], id) == p->value;
    }
    return 0;
}
static int body_die_if(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)out;
    return pred(c, st, id) ? 0 : 1;
}
static int body_add_self_edge(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)st; (void)out;
    const vh_test_params* p = (const vh_test_params*)c->s->params;
    ctx_add_edge(c, c->d->role_edge[0], id, id, p->edge_state);
    return 1;
}
/* add_edges!(sim, id, edges(sim, id, ET)) test/edges.jl:357-359 */
static int body_readd_edges(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)st; (void)out;
    int et = c->d->role_edge[0];
    cont_t* k = rd_cont(c, et, id);
    if (!k) return 1;
    uint32_t n = k->n, bytes = c->s->m->e[et].bytes;
    /* edges() returns the container itself; copy before appending (the write
     * container differs from the read one whenever ET is in read) */
    uint64_t* fr = (uint64_t*)malloc(n * 8 + 8);
    uint8_t* sb = (uint8_t*)malloc((size_t)n * bytes + 1);
    memcpy(fr, k->from, n * 8);
    if (bytes) memcpy(sb, k->st, (size_t)n * bytes);
    for (uint32_t i = 0; i < n && !c->err; i++) ctx_add_edge(c, et, fr[i], id, sb + (size_t)i * bytes);
    free(fr); free(sb);
    return 1;
}
/* test/independent.jl:55-68 and :119-129 */
static int body_neighbor_foo_edges(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)out;
    const vh_test_params* p = (const vh_test_params*)c->s->params;
    int64_t foo = rd_i64(st);
    if (foo == p->value) return 0;
    cont_t* k = rd_cont(c, c->d->role_edge[0], id);
    if (k) {
        uint32_t n = k->n;
        uint64_t* nb = (uint64_t*)malloc(n * 8 + 8);
        memcpy(nb, k->from, n * 8);
        uint8_t es[8];
        wr_i64(es, foo);
        for (uint32_t i = 0; i < n && !c->err; i++) {
            ctx_add_edge(c, c->d->role_edge[1], nb[i], id, es);
            ctx_add_edge(c, c->d->role_edge[1], id, nb[i], es);
        }
        free(nb);
    }
    return 1;
}
/* test/independent.jl:165-176: `value` births of T(foo + 10j); with cmp 0
 * each newborn is linked both ways by an edge carrying its id, with cmp 1
 * only newborn -> caller (test/addexisting.jl:41-44) */
static int body_spawn(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)out;
    const vh_test_params* p = (const vh_test_params*)c->s->params;
    int at = c->d->role_agent[0], et = c->d->role_edge[0];
    uint32_t bytes = (at >= 1 && at <= c->s->m->na) ? c->s->m->a[at].bytes : 0;
    uint32_t cbytes = c->s->m->a[type_nr(id)].bytes;
    uint8_t ns[64];
    for (int64_t j = 1; j <= p->value && !c->err; j++) {
        memset(ns, 0, sizeof ns);
        memcpy(ns, st, bytes < cbytes ? bytes : cbytes);
        if (bytes >= 8 && cbytes >= 8) wr_i64(ns, rd_i64(st) + 10 * j);
        uint64_t nid = ctx_add_agent(c, at, ns);
        if (et >= 0 && !c->err) {
            uint8_t es[8];
            wr_i64(es, (int64_t)nid);
            ctx_add_edge(c, et, nid, id, es);
            if (p->cmp == 0) ctx_add_edge(c, et, id, nid, es);
        }
    }
    return 1;
}
static int body_count_edges(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)st;
    wr_i64(out, (int64_t)num_edges_of(c, c->d->role_edge[0], id));
    return 1;
}

/* SIR contagion: state {UInt8 health, UInt16 days} (bytes 0 and 2). */
#define KEY_RECOVER (1ull << 40)
#define KEY_DEATH   (2ull << 40)
#define KEY_BIRTH   (3ull << 40)
#define KEY_PARTNER (4ull << 40)
static int body_sir(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    const vh_sir_params* p = (const vh_sir_params*)c->s->params;
    uint8_t h = st[0];
    uint16_t days;
    memcpy(&days, st + 2, 2);
    if (h == VH_S) {
        cont_t* k = rd_cont(c, c->d->role_edge[0], id);
        if (k) {
            for (uint32_t i = 0; i < k->n; i++) {
                const uint8_t* ns = agentstate_of(c, k->from[i]);
                if (!ns) return 1;
                if (ns[0] == VH_I && u01(or_hash4(p->seed, id, c->step, (uint64_t)i + 1)) < p->p_inf) {
                    out[0] = VH_I; days = 0; memcpy(out + 2, &days, 2);
                    return 1;
                }
            }
        }
    } else if (h == VH_I) {
        days = (uint16_t)(days + 1);
        out[0] = days >= p->recovery ? VH_R : VH_I;
        memcpy(out + 2, &days, 2);
    }
    return 1;
}

/* Weighted bounded confidence.  Canonical summation order (shared with the
 * device): degree <= 64 sequential in insertion order; degree > 64: lane l
 * of 64 sums positions k = l, l+64, ... in order, then an xor butterfly
 * p[l] = p[l] + p[l ^ o] for o = 32, 16, 8, 4, 2, 1; the result is lane 0. */
void or_canon_sum_f64(const double* nt, const double* dt, uint32_t d, double* num, double* den) {
    if (d <= 64) {
        double a = 0.0, b = 0.0;
        for (uint32_t k = 0; k < d; k++) { a = a + nt[k]; b = b + dt[k]; }
        *num = a; *den = b;
        return;
    }
    double pa[64], pb[64];
    for (int l = 0; l < 64; l++) {
        double a = 0.0, b = 0.0;
        for (uint32_t k = (uint32_t)l; k < d; k += 64) { a = a + nt[k]; b = b + dt[k]; }
        pa[l] = a; pb[l] = b;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        double qa[64], qb[64];
        for (int l = 0; l < 64; l++) { qa[l] = pa[l] + pa[l ^ o]; qb[l] = pb[l] + pb[l ^ o]; }
        memcpy(pa, qa, sizeof pa); memcpy(pb, qb, sizeof pb);
    }
    *num = pa[0]; *den = pb[0];
}
static void canon_sum_f32(const float* nt, const float* dt, uint32_t d, float* num, float* den) {
    if (d <= 64) {
        float a = 0.0f, b = 0.0f;
        for (uint32_t k = 0; k < d; k++) { a = a + nt[k]; b = b + dt[k]; }
        *num = a; *den = b;
        return;
    }
    float pa[64], pb[64];
    for (int l = 0; l < 64; l++) {
        float a = 0.0f, b = 0.0f;
        for (uint32_t k = (uint32_t)l; k < d; k += 64) { a = a + nt[k]; b = b + dt[k]; }
        pa[l] = a; pb[l] = b;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        float qa[64], qb[64];
        for (int l = 0; l < 64; l++) { qa[l] = pa[l] + pa[l ^ o]; qb[l] = pb[l] + pb[l ^ o]; }
        memcpy(pa, qa, sizeof pa); memcpy(pb, qb, sizeof pb);
    }
    *num = pa[0]; *den = pb[0];
}

/* x_i <- sum_{j: |x_j - x_i| <= eps} w_ij x_j / sum w_ij; empty set keeps x_i.
 * Rejected neighbours add +0.0 terms, which leaves both sums bit-identical to
 * skipping them (the sums start at +0.0 and all terms are >= 0 or exact). */
static int body_opinion_f64(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    const vh_opinion_params* p = (const vh_opinion_params*)c->s->params;
    double xi;
    memcpy(&xi, st, 8);
    int et = c->d->role_edge[0];
    cont_t* k = rd_cont(c, et, id);
    double res = xi;
    if (k && k->n) {
        uint32_t d = k->n;
        double* nt = (double*)malloc(d * sizeof(double));
        double* dt = (double*)malloc(d * sizeof(double));
        for (uint32_t i = 0; i < d; i++) {
            const uint8_t* ns = agentstate_of(c, k->from[i]);
            if (!ns) { free(nt); free(dt); return 1; }
            double xj; float w;
            memcpy(&xj, ns, 8);
            memcpy(&w, k->st + (size_t)i * 4, 4);
            int acc = fabs(xj - xi) <= p->eps;
            nt[i] = acc ? (double)w * xj : 0.0;
            dt[i] = acc ? (double)w : 0.0;
        }
        double num, den;
        or_canon_sum_f64(nt, dt, d, &num, &den);
        if (den > 0.0) res = num / den;
        free(nt); free(dt);
    }
    memcpy(out, &res, 8);
    return 1;
}
static int body_opinion_f32(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    const vh_opinion_params* p = (const vh_opinion_params*)c->s->params;
    float xi, eps = (float)p->eps;
    memcpy(&xi, st, 4);
    int et = c->d->role_edge[0];
    cont_t* k = rd_cont(c, et, id);
    float res = xi;
    if (k && k->n) {
        uint32_t d = k->n;
        float* nt = (float*)malloc(d * sizeof(float));
        float* dt = (float*)malloc(d * sizeof(float));
        for (uint32_t i = 0; i < d; i++) {
            const uint8_t* ns = agentstate_of(c, k->from[i]);
            if (!ns) { free(nt); free(dt); return 1; }
            float xj, w;
            memcpy(&xj, ns, 4);
            memcpy(&w, k->st + (size_t)i * 4, 4);
            int acc = fabsf(xj - xi) <= eps;
            nt[i] = acc ? w * xj : 0.0f;
            dt[i] = acc ? w : 0.0f;
        }
        float num, den;
        canon_sum_f32(nt, dt, d, &num, &den);
        if (den > 0.0f) res = num / den;
        free(nt); free(dt);
    }
    memcpy(out, &res, 4);
    return 1;
}

/* Dynamic contact network: state {UInt8 health, Float32 susceptibility}
 * (bytes 0 and 4).  Previous-step contacts (in-edges) drive infection; a
 * death (u < p_death) returns nothing and adds one newborn; survivors draw
 * n_contacts partners uniformly over the live slots of the read state
 * (rejecting dead slots and themselves, max_redraw draws each) and add the
 * edge partner -> self. */
static int body_contact(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    const vh_contact_params* p = (const vh_contact_params*)c->s->params;
    int t = type_nr(id);
    uint8_t h = st[0];
    float susc;
    memcpy(&susc, st + 4, 4);
    if (h == VH_S) {
        cont_t* k = rd_cont(c, c->d->role_edge[0], id);
        if (k) {
            double pinf = p->beta * (double)susc;
            for (uint32_t i = 0; i < k->n; i++) {
                const uint8_t* ns = agentstate_of(c, k->from[i]);
                if (!ns) return 1;
                if (ns[0] == VH_I && u01(or_hash4(p->seed, id, c->step, (uint64_t)i + 1)) < pinf) {
                    h = VH_I;
                    break;
                }
            }
        }
    } else if (h == VH_I) {
        if (u01(or_hash4(p->seed, id, c->step, KEY_RECOVER)) < p->gamma) h = VH_R;
    }
    if (u01(or_hash4(p->seed, id, c->step, KEY_DEATH)) < p->p_death) {
        uint8_t nb[8] = {0};
        float ns = 0.5f + 0.5f * u01f(or_hash4(p->seed, id, c->step, KEY_BIRTH));
        nb[0] = VH_S;
        memcpy(nb + 4, &ns, 4);
        ctx_add_agent(c, t, nb);
        return 0;
    }
    agents_t* A = &c->s->a[t];
    uint64_t S = A->read.len, me = agent_nr(id);
    for (uint32_t cc = 0; cc < p->n_contacts; cc++) {
        for (uint32_t r = 0; r < p->max_redraw; r++) {
            uint64_t hh = or_hash4(p->seed, id, c->step, KEY_PARTNER + (uint64_t)cc * p->max_redraw + r);
            uint64_t j = 1 + (((hh >> 32) * S) >> 32);
            if (j == me || A->read.died[j - 1]) continue;
            ctx_add_edge(c, c->d->role_edge[0], agent_id((uint64_t)t, j), id, NULL);
            break;
        }
    }
    out[0] = h;
    memcpy(out + 4, &susc, 4);
    return 1;
}

/* diffuse / diffuse3D test/raster.jl:38-41,143-146: the Bool field becomes
 * a.active || mapreduce(n -> n.active, |, neighborstates(sim, id, E)) */
static int body_diffuse(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    const vh_raster_params* p = (const vh_raster_params*)c->s->params;
    uint32_t f = p->field_offset;
    uint8_t v = st[f] ? 1 : 0;
    if (!v) {
        cont_t* k = rd_cont(c, c->d->role_edge[0], id);
        if (k)
            for (uint32_t i = 0; i < k->n; i++) {
                const uint8_t* ns = agentstate_of(c, k->from[i]);
                if (!ns) return 1;
                v |= ns[f] ? 1 : 0;
            }
    }
    out[f] = v;
    return 1;
}
/* value_on_pos test/raster.jl:279-285 with move_to! src/Raster.jl:437-463
 * (distance 0): v = first(neighborstates_flexible(sim, id, E)).ids_sum, then
 * add_edge!(cell(v,...,v), id) and add_edge!(id, cell(v,...,v)) */
static int body_move_to(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)st;
    const vh_raster_params* p = (const vh_raster_params*)c->s->params;
    int et = c->d->role_edge[0];
    cont_t* k = rd_cont(c, et, id);
    if (!k || k->n == 0) { c->err = fail(VH_ERR_ASSERT, "first(nothing): no neighbour"); return 1; }
    const uint8_t* ns = agentstate_of(c, k->from[0]);
    if (!ns) return 1;
    int64_t v = rd_i64(ns);
    uint64_t lin = 0, stride = 1;
    for (uint32_t d = 0; d < p->ndims && d < 4; d++) {
        if (v < 1 || v > p->dims[d]) { c->err = fail(VH_ERR_ASSERT, "raster position out of bounds"); return 1; }
        lin += (uint64_t)(v - 1) * stride;
        stride *= (uint64_t)p->dims[d];
    }
    uint64_t cell = agent_id(p->cell_type, p->first_nr + lin);
    if (p->flags & 1u) ctx_add_edge(c, et, cell, id, p->edge_state);
    if (p->flags & 2u) ctx_add_edge(c, et, id, cell, p->edge_state);
    wr_i64(out, v);
    return 1;
}

/* the remove_edges! bodies of test/mpi/test_edgetypes.jl:295-445 (enum
 * vh_remove_mode); idx = the Int64 at offset 0 of the caller's state */
static int nth_from(ctx_t* c, int et, uint64_t id, uint32_t n, uint64_t* out) {
    cont_t* k = rd_cont(c, et, id);
    if (!k || !k->present || k->n <= n) return 0;
    if (HAS(ETH(c->s, et), VH_HINT_IGNOREFROM)) {
        c->err = fail(VH_ERR_ASSERT, "neighborids is not defined for the hint combination");
        return 0;
    }
    *out = k->from[n];
    return 1;
}
static int body_remove_edges(ctx_t* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)out;
    const vh_test_params* p = (const vh_test_params*)c->s->params;
    int64_t idx = rd_i64(st);
    int et = c->d->role_edge[0], es = c->d->role_edge[1];
    uint64_t nb = 0, nb2 = 0;
    switch ((int)p->value) {
    case VH_RM_OWN:
        if (idx % 2 == 0) ctx_remove_edges(c, et, 0, id, 1);
        break;
    case VH_RM_NEIGHBOR:
        if (idx % 2 == 0) {
            if (!nth_from(c, et, id, 0, &nb)) { if (!c->err) c->err = fail(VH_ERR_ASSERT, "no neighbour"); }
            else ctx_remove_edges(c, et, 0, nb, 1);
        }
        break;
    case VH_RM_READD:
        if (!nth_from(c, et, id, 0, &nb)) { if (!c->err) c->err = fail(VH_ERR_ASSERT, "no neighbour"); break; }
        ctx_remove_edges(c, et, nb, id, 0);
        if (!c->err) ctx_add_edge(c, et, nb, id, p->edge_state);
        break;
    case VH_RM_FROM_TWO: {
        int single = HAS(ETH(c->s, et), VH_HINT_SINGLEEDGE);
        if (!nth_from(c, et, id, 0, &nb) || (!single && !nth_from(c, et, id, 1, &nb2))) {
            if (!c->err) c->err = fail(VH_ERR_ASSERT, "no neighbour");
            break;
        }
        ctx_remove_edges(c, et, nb, id, 0);
        if (!single && !c->err) ctx_remove_edges(c, et, nb2, id, 0);
        break;
    }
    case VH_RM_TO_NEIGHBOR:
        if (idx % 2 == 0) {
            if (!nth_from(c, et, id, 0, &nb)) { if (!c->err) c->err = fail(VH_ERR_ASSERT, "no neighbour"); }
            else ctx_remove_edges(c, et, id, nb, 0);
        }
        break;
    case VH_RM_SEND_ID:
        if (!nth_from(c, et, id, 0, &nb)) { if (!c->err) c->err = fail(VH_ERR_ASSERT, "no neighbour"); }
        else ctx_add_edge(c, es, id, nb, p->edge_state);
        break;
    case VH_RM_TO_SENDID:
        if (idx < 100) {
            if (!nth_from(c, es, id, 0, &nb)) { if (!c->err) c->err = fail(VH_ERR_ASSERT, "no neighbour"); }
            else ctx_remove_edges(c, et, id, nb, 0);
        }
        break;
    case VH_RM_ZERO_FROM:
        ctx_remove_edges(c, et, 0, id, 0);
        break;
    }
    return 1;
}

/* ------------------------------------------------ caller-supplied bodies
 * The checker's counterpart of vh_register_body: a C restatement of a
 * run-time body, registered as a function over the accessors below. */
#define OR_MAX_USER 64
static body_fn g_user[OR_MAX_USER];
static uint32_t g_nuser;

int or_register_body(or_body_fn fn, uint32_t flags, uint32_t* id) {
    (void)flags;   /* the checker applies add_edge!/add_agent! as they come */
    if (!fn || !id) return fail(VH_ERR_ARG, "null argument");
    for (uint32_t k = 0; k < g_nuser; k++)
        if (g_user[k] == (body_fn)fn) { *id = VH_TR_USER + k; return VH_OK; }
    if (g_nuser >= OR_MAX_USER) return fail(VH_ERR_CAPACITY, "too many registered bodies");
    g_user[g_nuser] = (body_fn)fn;
    *id = VH_TR_USER + g_nuser++;
    return VH_OK;
}
const void* or_ctx_params(or_ctx* c) { return ((ctx_t*)c)->s->params; }
int or_ctx_role_edge(or_ctx* c, int k) { return ((ctx_t*)c)->d->role_edge[k]; }
int or_ctx_role_agent(or_ctx* c, int k) { return ((ctx_t*)c)->d->role_agent[k]; }
uint32_t or_ctx_in_edges(or_ctx* cc, int et, uint64_t id, const uint64_t** from, const uint8_t** states) {
    ctx_t* c = (ctx_t*)cc;
    cont_t* k = rd_cont(c, et, id);
    if (!k) return 0;
    if (from) *from = k->from;
    if (states) *states = k->st;
    return k->n;
}
const uint8_t* or_ctx_agentstate(or_ctx* c, uint64_t id) { return agentstate_of((ctx_t*)c, id); }
void or_ctx_add_edge(or_ctx* c, int et, uint64_t from, uint64_t to, const void* st) {
    ctx_add_edge((ctx_t*)c, et, from, to, (const uint8_t*)st);
}
uint64_t or_ctx_add_agent(or_ctx* c, int t, const void* st) {
    return ctx_add_agent((ctx_t*)c, t, (const uint8_t*)st);
}
void or_ctx_remove_edges(or_ctx* c, int et, uint64_t from, uint64_t to, int all) {
    ctx_remove_edges((ctx_t*)c, et, from, to, all);
}
int or_ctx_failed(or_ctx* c) { return ((ctx_t*)c)->err != 0; }

static body_fn body_of(uint32_t tr) {
    if (tr >= VH_TR_USER && tr - VH_TR_USER < g_nuser) return g_user[tr - VH_TR_USER];
    switch (tr) {
    case VH_TR_NOOP: return body_noop;
    case VH_TR_SUM_NEIGHBORS: return body_sum_neighbors;
    case VH_TR_DIE_IF: return body_die_if;
    case VH_TR_ADD_SELF_EDGE: return body_add_self_edge;
    case VH_TR_READD_EDGES: return body_readd_edges;
    case VH_TR_NEIGHBOR_FOO_EDGES: return body_neighbor_foo_edges;
    case VH_TR_SPAWN: return body_spawn;
    case VH_TR_COUNT_EDGES: return body_count_edges;
    case VH_TR_DIFFUSE: return body_diffuse;
    case VH_TR_MOVE_TO: return body_move_to;
    case VH_TR_REMOVE_EDGES: return body_remove_edges;
    case VH_TR_SIR: return body_sir;
    case VH_TR_OPINION_F64: return body_opinion_f64;
    case VH_TR_OPINION_F32: return body_opinion_f32;
    case VH_TR_CONTACT: return body_contact;
    }
    return NULL;
}

/* ------------------------------------------------------------ apply! */
/* transition_with_write! src/AgentMethods.jl:159-181 */
static int write_result(or_sim* s, int t, uint64_t idx, int keep, const uint8_t* st) {
    agents_t* A = &s->a[t];
    uint32_t bytes = s->m->a[t].bytes;
    int indep = HAS(s->m->a[t].hints, VH_HINT_INDEPENDENT);
    if (!mortal_t(s, t)) {
        if (!keep) return fail(VH_ERR_ASSERT, "You can not return `nothing` for immortal agents");
    } else if (!keep) {
        u64_push(&A->write.reuse, idx);
        A->write.died[idx - 1] = 1;
        return VH_OK;
    }
    if (indep) memcpy(A->read.state + (idx - 1) * bytes, st, bytes);
    else memcpy(A->write.state + (idx - 1) * bytes, st, bytes);
    return VH_OK;
}

static int in_mask(uint32_t mask, int i) { return (mask >> i) & 1u; }

/* param(sim, name) src/Simulation.jl:587, set_param! :602-625 (asserts
 * after finish_init!), get_global / set_global! src/Global.jl:10-59, as
 * numbered slots holding one scalar each */
static uint32_t dt_bytes(int dt) {
    switch (dt) {
    case VH_DT_I64: case VH_DT_F64: case VH_DT_U64: return 8;
    case VH_DT_F32: case VH_DT_U32: case VH_DT_I32: return 4;
    case VH_DT_U16: return 2;
    case VH_DT_BOOL: case VH_DT_U8: return 1;
    default: return 0;
    }
}
int or_sim_set_value(or_sim* s, int kind, uint32_t slot, int dtype, const void* value) {
    if (kind != VH_VALUE_PARAM && kind != VH_VALUE_GLOBAL) return fail(VH_ERR_ARG, "kind");
    if (slot >= VH_MAX_VALUES || !dt_bytes(dtype) || !value) return fail(VH_ERR_ARG, "slot/dtype");
    if (kind == VH_VALUE_PARAM && s->initialized)
        return fail(VH_ERR_ASSERT, "You can call `set_param!` only before `finish_init!`.");
    uint64_t raw = 0;
    memcpy(&raw, value, dt_bytes(dtype));
    s->vals[kind][slot] = raw;
    s->vdt[kind][slot] = (uint8_t)dtype;
    s->vset[kind] |= 1u << slot;
    return VH_OK;
}
int or_sim_get_value(or_sim* s, int kind, uint32_t slot, int* dtype, void* value) {
    if ((kind != VH_VALUE_PARAM && kind != VH_VALUE_GLOBAL) || slot >= VH_MAX_VALUES ||
        !((s->vset[kind] >> slot) & 1u))
        return fail(VH_ERR_ARG, "value slot not set");
    if (dtype) *dtype = s->vdt[kind][slot];
    memcpy(value, &s->vals[kind][slot], 8);
    return VH_OK;
}
/* Julia's convert of a stored scalar to a field type: any number to a float
 * type; to an integer type only an integral value in its range (else an
 * InexactError); Bool takes 0 and 1. The value is read as long double, which
 * holds every Int64/UInt64 and Float64 exactly (x86 80-bit extended). */
static int convert_to(int from, uint64_t raw, int to, uint8_t* out) {
    long double v;
    switch (from) {
    case VH_DT_F64: { double x; memcpy(&x, &raw, 8); v = x; break; }
    case VH_DT_F32: { float x; memcpy(&x, &raw, 4); v = x; break; }
    case VH_DT_I64: { int64_t x; memcpy(&x, &raw, 8); v = (long double)x; break; }
    case VH_DT_I32: { int32_t x; memcpy(&x, &raw, 4); v = x; break; }
    case VH_DT_U64: { uint64_t x; memcpy(&x, &raw, 8); v = (long double)x; break; }
    case VH_DT_U32: { uint32_t x; memcpy(&x, &raw, 4); v = x; break; }
    case VH_DT_U16: { uint16_t x; memcpy(&x, &raw, 2); v = x; break; }
    case VH_DT_U8: case VH_DT_BOOL: { uint8_t x; memcpy(&x, &raw, 1); v = x; break; }
    default: return 0;
    }
    if (to == VH_DT_F64) { double x = (double)v; memcpy(out, &x, 8); return 1; }
    if (to == VH_DT_F32) { float x = (float)v; memcpy(out, &x, 4); return 1; }
    if (v != floorl(v)) return 0;
    switch (to) {
    case VH_DT_I64: if (v < -9223372036854775808.0L || v > 9223372036854775807.0L) return 0;
                    { int64_t x = (int64_t)v; memcpy(out, &x, 8); } return 1;
    case VH_DT_U64: if (v < 0 || v > 18446744073709551615.0L) return 0;
                    { uint64_t x = (uint64_t)v; memcpy(out, &x, 8); } return 1;
    case VH_DT_I32: if (v < -2147483648.0L || v > 2147483647.0L) return 0;
                    { int32_t x = (int32_t)v; memcpy(out, &x, 4); } return 1;
    case VH_DT_U32: if (v < 0 || v > 4294967295.0L) return 0;
                    { uint32_t x = (uint32_t)v; memcpy(out, &x, 4); } return 1;
    case VH_DT_U16: if (v < 0 || v > 65535.0L) return 0;
                    { uint16_t x = (uint16_t)v; memcpy(out, &x, 2); } return 1;
    case VH_DT_U8: if (v < 0 || v > 255.0L) return 0;
                   { uint8_t x = (uint8_t)v; memcpy(out, &x, 1); } return 1;
    case VH_DT_BOOL: if (v < 0 || v > 1.0L) return 0;
                     { uint8_t x = (uint8_t)v; memcpy(out, &x, 1); } return 1;
    default: return 0;
    }
}

static int or_apply_resolved(or_sim* s, const vh_apply_desc* d);

/* apply! src/Simulation.jl:720-821; blob fields bound to param/global
 * values are filled in first (the values are constant during a transition) */
int or_apply(or_sim* s, const vh_apply_desc* d) {
    if (!d->n_bind) return or_apply_resolved(s, d);
    uint8_t blob[256];
    vh_apply_desc r = *d;
    if (d->params_bytes > sizeof blob || !d->bind) return fail(VH_ERR_ARG, "params/bind");
    memset(blob, 0, sizeof blob);
    if (d->params_bytes) memcpy(blob, d->params, d->params_bytes);
    for (uint32_t k = 0; k < d->n_bind; k++) {
        const vh_value_bind* b = &d->bind[k];
        uint32_t sz = dt_bytes(b->dtype);
        if ((b->kind != VH_VALUE_PARAM && b->kind != VH_VALUE_GLOBAL) || b->slot >= VH_MAX_VALUES || !sz ||
            (uint64_t)b->offset + sz > d->params_bytes)
            return fail(VH_ERR_ARG, "invalid value binding");
        if (!((s->vset[b->kind] >> b->slot) & 1u)) return fail(VH_ERR_ASSERT, "value not set");
        if (!convert_to(s->vdt[b->kind][b->slot], s->vals[b->kind][b->slot], b->dtype, blob + b->offset))
            return fail(VH_ERR_ASSERT, "InexactError");
    }
    r.params = blob;
    r.n_bind = 0;
    r.bind = NULL;
    return or_apply_resolved(s, &r);
}

static int or_apply_resolved(or_sim* s, const vh_apply_desc* d) {
    or_model* m = s->m;
    if (!s->initialized) return fail(VH_ERR_ASSERT, "You must call finish_init! before apply!");
    if (d->with_edge >= 0 && HAS(m->e[d->with_edge].hints, VH_HINT_SINGLETYPE))
        return fail(VH_ERR_ASSERT, "The `with_edge` keyword can only be used for edgetypes "
                                   "without the :SingleType hint.");
    body_fn body = body_of(d->transition);
    if (!body) return fail(VH_ERR_UNREGISTERED, "unregistered transition %u", d->transition);
    if (d->params_bytes > sizeof s->params) return fail(VH_ERR_ARG, "params too large");
    memset(s->params, 0, sizeof s->params);
    if (d->params_bytes) memcpy(s->params, d->params, d->params_bytes);
    s->intransition = 1;
    uint32_t callmask = 0;
    for (uint32_t i = 0; i < d->n_call; i++) callmask |= 1u << d->call[i];
    if (callmask & d->add_existing_agents) return fail(VH_ERR_ASSERT, "call type can not be element of `add_existing`");
    if ((d->add_existing_agents & ~d->write_agents) || (d->add_existing_edges & ~d->write_edges))
        return fail(VH_ERR_ASSERT, "a type is in `add_existing` but not in `write`");
    s->d = d;
    /* prepare_read! */
    for (int et = 0; et < m->ne; et++) if (in_mask(d->read_edges, et)) s->e[et].readable = 1;
    for (int t = 1; t <= m->na; t++) if (in_mask(d->read_agents, t)) s->a[t].prepared = 1;
    for (int et = 0; et < m->ne; et++)
        if (in_mask(d->write_edges, et) && !in_mask(d->add_existing_edges, et)) aot_clear(&s->e[et]);
    /* prepare_write! */
    for (int t = 1; t <= m->na; t++) {
        if (!in_mask(d->write_agents, t)) continue;
        int ae = in_mask(callmask, t) || in_mask(d->add_existing_agents, t);
        if (!mortal_t(s, t) && !ae)
            return fail(VH_ERR_ASSERT, "You can add the type to the `write` argument only when it is "
                                       "also included in the `add_existing` argument, as it has the "
                                       "hint :Immortal.");
        agents_t* A = &s->a[t];
        if (!ae) {
            rw_free(&A->read); rw_free(&A->write);
            A->nextid = 1;
        }
        A->writeable = 1;
    }
    for (int et = 0; et < m->ne; et++) {
        if (!in_mask(d->write_edges, et)) continue;
        edges_t* X = &s->e[et];
        if (in_mask(d->add_existing_edges, et)) {
            X->write = in_mask(d->read_edges, et) ? field_copy(X->read, m->e[et].bytes) : X->read;
        } else {
            X->write = field_new();
        }
        X->writeable = 1;
        X->add_existing = in_mask(d->add_existing_edges, et);
    }
    /* the agent loop, call types in order */
    ctx_t c = {s, d, 0, s->num_transitions, 0};
    uint8_t out[256];
    int rc = VH_OK;
    for (uint32_t ci = 0; ci < d->n_call && rc == VH_OK; ci++) {
        int t = d->call[ci];
        if (t < 1 || t > m->na) { rc = fail(VH_ERR_UNREGISTERED, "unregistered call type"); break; }
        c.call = t;
        agents_t* A = &s->a[t];
        uint32_t bytes = m->a[t].bytes;
        int wr = in_mask(d->write_agents, t), mortal = mortal_t(s, t);
        static const uint8_t zero[256];
        if (d->with_edge < 0) {
            /* transition_with_read! / _without_read! src/AgentMethods.jl:193-245 */
            uint64_t len = A->read.len;
            for (uint64_t idx = 1; idx <= len; idx++) {
                if (mortal && A->read.died[idx - 1]) continue;
                const uint8_t* st = bytes ? A->read.state + (idx - 1) * bytes : zero;
                memcpy(out, st, bytes);
                int keep = body(&c, st, agent_id((uint64_t)t, idx), out);
                if (c.err) { rc = c.err; break; }
                if (wr && (rc = write_result(s, t, idx, keep, out))) break;
            }
        } else {
            /* transition_with_read_with_edge! src/AgentMethods.jl:209-228 */
            field_t* f = s->e[d->with_edge].read;
            uint64_t nrows = f->nrows[t];
            for (uint64_t nr = 1; nr <= nrows; nr++) {
                if (!f->rows[t][nr - 1].present) continue;
                if (mortal && nr <= A->read.len && A->read.died[nr - 1]) {
                    rc = fail(VH_ERR_ASSERT, "with_edge called a removed agent"); break;
                }
                const uint8_t* st = bytes ? A->read.state + (nr - 1) * bytes : zero;
                memcpy(out, st, bytes);
                int keep = body(&c, st, agent_id((uint64_t)t, nr), out);
                if (c.err) { rc = c.err; break; }
                if (wr && (rc = write_result(s, t, nr, keep, out))) break;
            }
        }
    }
    if (rc != VH_OK) return rc;   /* like an AssertionError: sim left mid-transition */
    /* finish_read! */
    for (int et = 0; et < m->ne; et++) if (in_mask(d->read_edges, et)) s->e[et].readable = 0;
    for (int t = 1; t <= m->na; t++) if (in_mask(d->read_agents, t)) s->a[t].prepared = 0;
    /* finish_write! agents then edges */
    for (int t = 1; t <= m->na; t++) if (in_mask(d->write_agents, t)) finish_write_agents(s, t);
    for (int et = 0; et < m->ne; et++) if (in_mask(d->write_edges, et)) finish_write_edges(s, et);
    s->intransition = 0;
    s->num_transitions++;
    s->d = NULL;
    return VH_OK;
}

int or_num_transitions(or_sim* s, uint64_t* out) { *out = s->num_transitions; return VH_OK; }

/* read_snapshot! src/HDF5.jl:1594-1646 with _read_agents_restore! :801-850:
 * the died flags of the n slots of type t; the reuseable slots become the
 * died ones in ascending order (not the order they died in) */
int or_restore_died(or_sim* s, uint8_t t, const uint8_t* died, uint64_t n) {
    if (t < 1 || t > s->m->na) return fail(VH_ERR_UNREGISTERED, "unregistered agent type");
    if (!s->initialized) return fail(VH_ERR_ASSERT, "restore after finish_init! only");
    agents_t* A = &s->a[t];
    if (n != A->read.len) return fail(VH_ERR_ARG, "restore: %llu slots, the type has %llu",
                                      (unsigned long long)n, (unsigned long long)A->read.len);
    if (!mortal_t(s, t)) {
        for (uint64_t i = 0; i < n; i++)
            if (died[i]) return fail(VH_ERR_ASSERT, "restore: a died agent of an :Immortal type");
        return VH_OK;
    }
    /* the reference reads the flags into write.died and publishes them
     * (finish_write!), so both buffers hold them */
    if (A->write.len < n) rw_resize(&A->write, n, s->m->a[t].bytes, 1);
    A->read.reuse.n = 0;
    for (uint64_t i = 0; i < n; i++) {
        A->read.died[i] = died[i] ? 1 : 0;
        A->write.died[i] = A->read.died[i];
        if (died[i]) u64_push(&A->read.reuse, i + 1);
    }
    return VH_OK;
}
int or_set_num_transitions(or_sim* s, uint64_t n) {
    s->num_transitions = n;
    return VH_OK;
}

/* ------------------------------------------------------------ queries */
/* num_agents src/Agent.jl:324-343 */
int or_num_agents(or_sim* s, uint8_t t, uint64_t* out) {
    if (t < 1 || t > s->m->na) return fail(VH_ERR_UNREGISTERED, "unregistered agent type");
    agents_t* A = &s->a[t];
    if (!mortal_t(s, t)) { *out = A->nextid - 1; return VH_OK; }
    rw_t* r = s->initialized ? &A->read : &A->write;
    uint64_t n = 0;
    for (uint64_t i = 0; i < r->len; i++) n += !r->died[i];
    *out = n;
    return VH_OK;
}
/* _num_edges src/EdgeMethods.jl:931-969 */
int or_num_edges(or_sim* s, uint8_t et, uint64_t* out) {
    if (et >= s->m->ne) return fail(VH_ERR_UNREGISTERED, "unregistered edge type");
    field_t* f = s->initialized ? s->e[et].read : s->e[et].write;
    uint64_t n = 0;
    for (int t = 0; t <= MAXA; t++)
        for (uint64_t i = 0; i < f->nrows[t]; i++)
            if (f->rows[t][i].present) n += f->rows[t][i].n;
    *out = n;
    return VH_OK;
}
int or_read_agents(or_sim* s, uint8_t t, void* states, uint8_t* died, uint64_t cap, uint64_t* nslots) {
    if (t < 1 || t > s->m->na) return fail(VH_ERR_UNREGISTERED, "unregistered agent type");
    rw_t* r = s->initialized ? &s->a[t].read : &s->a[t].write;
    uint64_t len = s->a[t].nextid - 1;
    if (len > r->len) len = r->len;
    *nslots = len;
    if (cap < len) return fail(VH_ERR_CAPACITY, "capacity");
    uint32_t bytes = s->m->a[t].bytes;
    if (states && bytes) memcpy(states, r->state, len * bytes);
    if (died) {
        if (mortal_t(s, t)) memcpy(died, r->died, len);
        else memset(died, 0, len);
    }
    return VH_OK;
}
/* edges_iterator order src/EdgeIterator.jl:45-106 (keys ascending) */
int or_read_edges(or_sim* s, uint8_t et, uint64_t* to, uint64_t* from, void* states,
                  uint64_t cap, uint64_t* n) {
    if (et >= s->m->ne) return fail(VH_ERR_UNREGISTERED, "unregistered edge type");
    field_t* f = s->e[et].read;
    uint32_t h = ETH(s, et), bytes = s->m->e[et].bytes;
    uint64_t k = 0;
    for (int t = 0; t <= MAXA; t++)
        for (uint64_t i = 0; i < f->nrows[t]; i++) {
            cont_t* c = &f->rows[t][i];
            if (!c->present) continue;
            for (uint32_t j = 0; j < c->n; j++, k++) {
                if (k >= cap) continue;
                if (to) to[k] = agent_id((uint64_t)t, i + 1);
                if (from) from[k] = HAS(h, VH_HINT_IGNOREFROM) ? 0 : c->from[j];
                if (states && bytes) memcpy((uint8_t*)states + k * bytes, c->st + (size_t)j * bytes, bytes);
            }
        }
    *n = k;
    return k > cap ? fail(VH_ERR_CAPACITY, "capacity") : VH_OK;
}
int or_edges_of(or_sim* s, uint8_t et, uint64_t to, uint64_t* from, void* states, uint64_t cap, uint64_t* n) {
    if (et >= s->m->ne) return fail(VH_ERR_UNREGISTERED, "unregistered edge type");
    cont_t* c = field_get(s->e[et].read, to);
    uint32_t h = ETH(s, et), bytes = s->m->e[et].bytes;
    *n = c ? c->n : 0;
    if (!c) return VH_OK;
    if (cap < c->n) return fail(VH_ERR_CAPACITY, "capacity");
    for (uint32_t j = 0; j < c->n; j++) {
        if (from) from[j] = HAS(h, VH_HINT_IGNOREFROM) ? 0 : c->from[j];
        if (states && bytes) memcpy((uint8_t*)states + (size_t)j * bytes, c->st + (size_t)j * bytes, bytes);
    }
    return VH_OK;
}
int or_agentstate(or_sim* s, uint64_t id, uint8_t t, void* out) {
    if (type_nr(id) != t) return fail(VH_ERR_ASSERT, "The id of the agent does not match the given type");
    rw_t* r = &s->a[t].read;
    uint64_t nr = agent_nr(id);
    if (nr == 0 || nr > r->len) return fail(VH_ERR_ASSERT, "agent nr out of range");
    if (mortal_t(s, t) && r->died[nr - 1]) return fail(VH_ERR_ASSERT, "agent has been removed");
    memcpy(out, r->state + (nr - 1) * s->m->a[t].bytes, s->m->a[t].bytes);
    return VH_OK;
}
int or_reuse_stack(or_sim* s, uint8_t t, uint64_t* out, uint64_t cap, uint64_t* n) {
    u64vec* r = &s->a[t].read.reuse;
    *n = r->n;
    if (cap < r->n) return fail(VH_ERR_CAPACITY, "capacity");
    memcpy(out, r->v, r->n * 8);
    return VH_OK;
}

/* ------------------------------------------------------------ mapreduce */
typedef union { int64_t i; double f; } acc_t;

static double field_f(const uint8_t* p, uint32_t dt) {
    switch (dt) {
    case VH_DT_F64: { double v; memcpy(&v, p, 8); return v; }
    case VH_DT_F32: { float v; memcpy(&v, p, 4); return v; }
    }
    return 0;
}
static int64_t field_i(const uint8_t* p, uint32_t dt) {
    switch (dt) {
    case VH_DT_I64: case VH_DT_U64: { int64_t v; memcpy(&v, p, 8); return v; }
    case VH_DT_I32: { int32_t v; memcpy(&v, p, 4); return v; }
    case VH_DT_U32: { uint32_t v; memcpy(&v, p, 4); return v; }
    case VH_DT_U16: { uint16_t v; memcpy(&v, p, 2); return v; }
    case VH_DT_U8: case VH_DT_BOOL: return p[0];
    case VH_DT_F64: return (int64_t)field_f(p, dt);
    case VH_DT_F32: return (int64_t)field_f(p, dt);
    }
    return 0;
}
static int is_fdt(int dt) { return dt == VH_DT_F64 || dt == VH_DT_F32; }
static acc_t map_val(const vh_map_desc* f, const uint8_t* st, int dtype) {
    acc_t a;
    if (f->kind == VH_MAP_ONE) {
        if (is_fdt(dtype)) a.f = 1.0; else a.i = 1;
        return a;
    }
    const uint8_t* p = st + f->offset;
    if (f->kind == VH_MAP_FIELD_EQ) {
        int eq = is_fdt(f->field_dtype) ? field_f(p, f->field_dtype) == (double)f->value
                                        : field_i(p, f->field_dtype) == f->value;
        if (is_fdt(dtype)) a.f = eq; else a.i = eq;
        return a;
    }
    if (is_fdt(dtype)) a.f = is_fdt(f->field_dtype) ? field_f(p, f->field_dtype) : (double)field_i(p, f->field_dtype);
    else a.i = field_i(p, f->field_dtype);
    return a;
}
/* Julia's max/min on floats: NaN wins, -0.0 < +0.0 */
static double jl_fmax(double a, double b) {
    if (a != a || b != b) return a != a ? a : b;
    if (a == b) return signbit(a) ? b : a;
    return a > b ? a : b;
}
static double jl_fmin(double a, double b) {
    if (a != a || b != b) return a != a ? a : b;
    if (a == b) return signbit(a) ? a : b;
    return a < b ? a : b;
}
static acc_t op_apply(int op, int dtype, acc_t x, acc_t y) {
    acc_t r;
    if (is_fdt(dtype)) {
        switch (op) {
        case VH_OP_ADD: r.f = x.f + y.f; break;
        case VH_OP_MUL: r.f = x.f * y.f; break;
        case VH_OP_MAX: r.f = jl_fmax(x.f, y.f); break;
        case VH_OP_MIN: r.f = jl_fmin(x.f, y.f); break;
        default: r.f = 0;
        }
        return r;
    }
    uint64_t a = (uint64_t)x.i, b = (uint64_t)y.i;
    switch (op) {
    case VH_OP_ADD: r.i = (int64_t)(a + b); break;
    case VH_OP_MUL: r.i = (int64_t)(a * b); break;
    case VH_OP_MAX: r.i = x.i > y.i ? x.i : y.i; break;
    case VH_OP_MIN: r.i = x.i < y.i ? x.i : y.i; break;
    case VH_OP_AND: r.i = x.i & y.i; break;
    case VH_OP_OR: r.i = x.i | y.i; break;
    default: r.i = 0;
    }
    if (dtype == VH_DT_BOOL) r.i = r.i != 0;
    return r;
}
/* val4empty src/Helpers.jl:44-81 */
static int val4empty(int op, int dtype, acc_t* out) {
    int f = is_fdt(dtype);
    switch (op) {
    case VH_OP_ADD: if (f) out->f = 0.0; else out->i = 0; return VH_OK;
    case VH_OP_MUL: if (f) out->f = 1.0; else out->i = 1; return VH_OK;
    case VH_OP_MAX:
        if (f) out->f = -INFINITY; else out->i = dtype == VH_DT_BOOL ? -1 : -INT64_MAX;
        return VH_OK;
    case VH_OP_MIN:
        if (f) out->f = INFINITY; else out->i = dtype == VH_DT_BOOL ? 1 : INT64_MAX;
        return VH_OK;
    case VH_OP_AND:
        if (f) return fail(VH_ERR_ASSERT, "The & operator is only supported for integer and boolean types");
        out->i = dtype == VH_DT_BOOL ? 1 : INT64_MAX;
        return VH_OK;
    case VH_OP_OR:
        if (f) return fail(VH_ERR_ASSERT, "The | operator is only supported for integer and boolean types");
        out->i = 0;
        return VH_OK;
    }
    return fail(VH_ERR_ASSERT, "Can not derive the init value for the operator.");
}
/* identity of each op for the padding of the device tree */
static acc_t op_identity(int op, int dtype) {
    acc_t a;
    int f = is_fdt(dtype);
    switch (op) {
    case VH_OP_ADD: case VH_OP_OR: if (f) a.f = 0.0; else a.i = 0; break;
    case VH_OP_MUL: if (f) a.f = 1.0; else a.i = 1; break;
    case VH_OP_MAX: if (f) a.f = -INFINITY; else a.i = INT64_MIN; break;
    case VH_OP_MIN: if (f) a.f = INFINITY; else a.i = INT64_MAX; break;
    case VH_OP_AND: a.i = dtype == VH_DT_BOOL ? 1 : -1; break;
    default: a.i = 0;
    }
    return a;
}
/* The device's fixed reduction tree: blocks of 2048 values; thread t of 256
 * folds v[t + 256 j], j = 0..7, starting from the identity; then pairs
 * a[t] = a[t] op a[t + s] for s = 128 .. 1; block results are reduced the
 * same way until one value remains.  Missing slots are the identity. */
static acc_t tree_reduce(acc_t* v, uint64_t n, int op, int dtype) {
    acc_t id = op_identity(op, dtype);
    if (n == 0) return id;
    while (1) {
        uint64_t nb = (n + 2047) / 2048;
        acc_t* r = (acc_t*)malloc(nb * sizeof(acc_t));
        for (uint64_t b = 0; b < nb; b++) {
            acc_t a[256];
            for (int t = 0; t < 256; t++) {
                acc_t x = id;
                for (int j = 0; j < 8; j++) {
                    uint64_t e = b * 2048 + (uint64_t)t + 256u * (uint64_t)j;
                    if (e < n) x = op_apply(op, dtype, x, v[e]);
                }
                a[t] = x;
            }
            for (int st = 128; st >= 1; st >>= 1)
                for (int t = 0; t < st; t++) a[t] = op_apply(op, dtype, a[t], a[t + st]);
            r[b] = a[0];
        }
        if (nb == 1) { acc_t x = r[0]; free(r); return x; }
        memcpy(v, r, nb * sizeof(acc_t));
        free(r);
        n = nb;
    }
}

/* ---- computed maps (vh_mapreduce_expr): the postfix program of
 * include/vahana_hip.h evaluated per record with a tagged stack */
enum { EX_I = 0, EX_F = 1, EX_B = 2 };
typedef struct { int tag; int64_t i; double f; } exv_t;

static double exv_f(exv_t v) { return v.tag == EX_F ? v.f : (double)v.i; }
static int exv_true(exv_t v) { return v.tag == EX_F ? v.f != 0.0 : v.i != 0; }
static uint32_t ex_field_bytes(uint32_t dt) {
    switch (dt) {
    case VH_DT_I64: case VH_DT_F64: case VH_DT_U64: return 8;
    case VH_DT_F32: case VH_DT_U32: case VH_DT_I32: return 4;
    case VH_DT_U16: return 2;
    case VH_DT_BOOL: case VH_DT_U8: return 1;
    }
    return 0;
}
static int ex_check(const vh_expr_op* p, uint32_t n, uint32_t bytes) {
    int d = 0;
    if (!p || n == 0 || n > VH_EXPR_MAX_OPS) return fail(VH_ERR_ARG, "mapreduce program: 1 to 64 operations");
    for (uint32_t k = 0; k < n; k++) {
        int pop;
        switch (p[k].code) {
        case VH_EX_FIELD:
            if (!ex_field_bytes(p[k].dtype) || p[k].arg + ex_field_bytes(p[k].dtype) > bytes)
                return fail(VH_ERR_ARG, "mapreduce program: field outside the state");
            pop = 0; break;
        case VH_EX_CONST_I: case VH_EX_CONST_F: case VH_EX_CONST_B: pop = 0; break;
        case VH_EX_NEG: case VH_EX_ABS: case VH_EX_NOT: pop = 1; break;
        case VH_EX_SELECT: pop = 3; break;
        case VH_EX_ADD: case VH_EX_SUB: case VH_EX_MUL: case VH_EX_DIV: case VH_EX_MIN: case VH_EX_MAX:
        case VH_EX_EQ: case VH_EX_NE: case VH_EX_LT: case VH_EX_LE: case VH_EX_GT: case VH_EX_GE:
        case VH_EX_AND: case VH_EX_OR: pop = 2; break;
        default: return fail(VH_ERR_ARG, "mapreduce program: unknown operation");
        }
        if (d < pop) return fail(VH_ERR_ARG, "mapreduce program: missing operands");
        d += 1 - pop;
        if (d > VH_EXPR_MAX_DEPTH) return fail(VH_ERR_ARG, "mapreduce program: stack too deep");
    }
    if (d != 1) return fail(VH_ERR_ARG, "mapreduce program: not one value");
    return VH_OK;
}
static exv_t ex_eval(const vh_expr_op* p, uint32_t n, const uint8_t* st) {
    exv_t stk[VH_EXPR_MAX_DEPTH];
    int sp = 0;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t c = p[k].code;
        exv_t r = {EX_I, 0, 0.0};
        if (c == VH_EX_FIELD) {
            const uint8_t* q = st + p[k].arg;
            if (is_fdt((int)p[k].dtype)) { r.tag = EX_F; r.f = field_f(q, p[k].dtype); }
            else if (p[k].dtype == VH_DT_BOOL) { r.tag = EX_B; r.i = q[0] != 0; }
            else { r.tag = EX_I; r.i = field_i(q, p[k].dtype); }
            stk[sp++] = r;
            continue;
        }
        if (c == VH_EX_CONST_I) { r.tag = EX_I; memcpy(&r.i, &p[k].arg, 8); stk[sp++] = r; continue; }
        if (c == VH_EX_CONST_F) { r.tag = EX_F; memcpy(&r.f, &p[k].arg, 8); stk[sp++] = r; continue; }
        if (c == VH_EX_CONST_B) { r.tag = EX_B; r.i = p[k].arg != 0; stk[sp++] = r; continue; }
        if (c == VH_EX_NEG || c == VH_EX_ABS || c == VH_EX_NOT) {
            exv_t a = stk[sp - 1];
            if (c == VH_EX_NOT) { r.tag = EX_B; r.i = !exv_true(a); }
            else if (a.tag == EX_F) { r.tag = EX_F; r.f = c == VH_EX_NEG ? -a.f : fabs(a.f); }
            else {
                uint64_t u = (uint64_t)a.i, neg = 0ull - u;
                r.tag = EX_I;
                r.i = (int64_t)(c == VH_EX_NEG ? neg : (a.i < 0 ? neg : u));
            }
            stk[sp - 1] = r;
            continue;
        }
        if (c == VH_EX_SELECT) {
            exv_t b = stk[sp - 1], a = stk[sp - 2], cc = stk[sp - 3];
            sp -= 2;
            exv_t x = exv_true(cc) ? a : b;
            if (a.tag == EX_F || b.tag == EX_F) { r.tag = EX_F; r.f = exv_f(x); }
            else { r.tag = (a.tag == EX_B && b.tag == EX_B) ? EX_B : EX_I; r.i = x.i; }
            stk[sp - 1] = r;
            continue;
        }
        {
            exv_t b = stk[sp - 1], a = stk[sp - 2];
            int fl = a.tag == EX_F || b.tag == EX_F;
            sp--;
            if (c >= VH_EX_EQ && c <= VH_EX_GE) {
                int t;
                if (fl) {
                    double x = exv_f(a), y = exv_f(b);
                    t = c == VH_EX_EQ ? x == y : c == VH_EX_NE ? x != y : c == VH_EX_LT ? x < y
                      : c == VH_EX_LE ? x <= y : c == VH_EX_GT ? x > y : x >= y;
                } else {
                    t = c == VH_EX_EQ ? a.i == b.i : c == VH_EX_NE ? a.i != b.i : c == VH_EX_LT ? a.i < b.i
                      : c == VH_EX_LE ? a.i <= b.i : c == VH_EX_GT ? a.i > b.i : a.i >= b.i;
                }
                r.tag = EX_B;
                r.i = t;
            } else if (c == VH_EX_AND || c == VH_EX_OR) {
                r.tag = EX_B;
                r.i = c == VH_EX_AND ? (exv_true(a) && exv_true(b)) : (exv_true(a) || exv_true(b));
            } else if (fl || c == VH_EX_DIV) {
                double x = exv_f(a), y = exv_f(b);
                r.tag = EX_F;
                switch (c) {
                case VH_EX_ADD: r.f = x + y; break;
                case VH_EX_SUB: r.f = x - y; break;
                case VH_EX_MUL: r.f = x * y; break;
                case VH_EX_DIV: r.f = x / y; break;
                case VH_EX_MIN: r.f = jl_fmin(x, y); break;
                case VH_EX_MAX: r.f = jl_fmax(x, y); break;
                }
            } else {
                uint64_t x = (uint64_t)a.i, y = (uint64_t)b.i;
                r.tag = EX_I;
                switch (c) {
                case VH_EX_ADD: r.i = (int64_t)(x + y); break;
                case VH_EX_SUB: r.i = (int64_t)(x - y); break;
                case VH_EX_MUL: r.i = (int64_t)(x * y); break;
                case VH_EX_MIN: r.i = a.i < b.i ? a.i : b.i; break;
                case VH_EX_MAX: r.i = a.i > b.i ? a.i : b.i; break;
                }
            }
            stk[sp - 1] = r;
        }
    }
    return stk[0];
}
/* the program's value in the reduction's type */
static acc_t ex_to_acc(exv_t v, int dtype) {
    acc_t a;
    if (is_fdt(dtype)) a.f = exv_f(v);
    else if (dtype == VH_DT_BOOL) a.i = exv_true(v);
    else if (v.tag != EX_F) a.i = v.i;
    else if (v.f != v.f) a.i = 0;                       /* NaN */
    else if (v.f >= 9223372036854775808.0) a.i = INT64_MAX;
    else if (v.f < -9223372036854775808.0) a.i = INT64_MIN;
    else a.i = (int64_t)v.f;
    return a;
}

/* mapreduce with a computed f: or_mapreduce's folds over the program's
 * values (order 0: the reference's sequential fold, 1: the device tree) */
int or_mapreduce_expr(or_sim* s, int kind, uint8_t t, const vh_expr_op* prog, uint32_t nops, int op, int dtype,
                      const void* init, void* out, int order) {
    if (s->intransition) return fail(VH_ERR_ASSERT, "You can not call mapreduce inside of a transition function.");
    acc_t acc = {0};
    if (init) { if (is_fdt(dtype)) memcpy(&acc.f, init, 8); else memcpy(&acc.i, init, 8); }
    else { int r = val4empty(op, dtype, &acc); if (r) return r; }
    acc_t* vals = NULL;
    uint64_t n = 0, cap = 0;
#define PUSHX(x) do { if (n == cap) { cap = cap ? cap * 2 : 1024; vals = (acc_t*)realloc(vals, cap * sizeof(acc_t)); } vals[n++] = (x); } while (0)
    if (kind == 0) {
        if (t < 1 || t > s->m->na) return fail(VH_ERR_UNREGISTERED, "unregistered agent type");
        agents_t* A = &s->a[t];
        uint32_t bytes = s->m->a[t].bytes;
        int r = ex_check(prog, nops, bytes);
        if (r) return r;
        int mortal = mortal_t(s, t);
        acc_t id = op_identity(op, dtype);
        for (uint64_t i = 1; i <= A->nextid - 1; i++) {
            if (mortal && A->read.died[i - 1]) { if (order) PUSHX(id); continue; }
            PUSHX(ex_to_acc(ex_eval(prog, nops, A->read.state + (i - 1) * bytes), dtype));
        }
    } else {
        if (t >= s->m->ne) return fail(VH_ERR_UNREGISTERED, "unregistered edge type");
        if (HAS(ETH(s, t), VH_HINT_STATELESS))
            return fail(VH_ERR_ASSERT, "mapreduce is not defined for the hint combination");
        uint32_t bytes = s->m->e[t].bytes;
        int r = ex_check(prog, nops, bytes);
        if (r) return r;
        field_t* fl = s->e[t].read;
        for (int at = 0; at <= MAXA; at++)
            for (uint64_t i = 0; i < fl->nrows[at]; i++) {
                cont_t* c = &fl->rows[at][i];
                if (!c->present) continue;
                for (uint32_t j = 0; j < c->n; j++)
                    PUSHX(ex_to_acc(ex_eval(prog, nops, c->st + (size_t)j * bytes), dtype));
            }
    }
    if (order == 0) {
        for (uint64_t i = 0; i < n; i++) acc = op_apply(op, dtype, vals[i], acc);
    } else {
        acc_t tot = tree_reduce(vals, n, op, dtype);
        acc = op_apply(op, dtype, tot, acc);
    }
    free(vals);
    if (is_fdt(dtype)) memcpy(out, &acc.f, 8);
    else memcpy(out, &acc.i, 8);
    return VH_OK;
#undef PUSHX
}

int or_mapreduce(or_sim* s, int kind, uint8_t t, const vh_map_desc* f, int op, int dtype,
                 const void* init, void* out, int order) {
    if (s->intransition) return fail(VH_ERR_ASSERT, "You can not call mapreduce inside of a transition function.");
    acc_t acc = {0};
    if (init) { if (is_fdt(dtype)) memcpy(&acc.f, init, 8); else memcpy(&acc.i, init, 8); }
    else { int r = val4empty(op, dtype, &acc); if (r) return r; }
    acc_t* vals = NULL;
    uint64_t n = 0, cap = 0;
#define PUSHV(x) do { if (n == cap) { cap = cap ? cap * 2 : 1024; vals = (acc_t*)realloc(vals, cap * sizeof(acc_t)); } vals[n++] = (x); } while (0)
    if (kind == 0) {
        if (t < 1 || t > s->m->na) return fail(VH_ERR_UNREGISTERED, "unregistered agent type");
        agents_t* A = &s->a[t];
        uint32_t bytes = s->m->a[t].bytes;
        int mortal = mortal_t(s, t);
        acc_t id = op_identity(op, dtype);
        for (uint64_t i = 1; i <= A->nextid - 1; i++) {
            if (mortal && A->read.died[i - 1]) { if (order) PUSHV(id); continue; }
            PUSHV(map_val(f, A->read.state + (i - 1) * bytes, dtype));
        }
    } else {
        if (t >= s->m->ne) return fail(VH_ERR_UNREGISTERED, "unregistered edge type");
        if (HAS(ETH(s, t), VH_HINT_STATELESS))
            return fail(VH_ERR_ASSERT, "mapreduce is not defined for the hint combination");
        field_t* fl = s->e[t].read;
        uint32_t bytes = s->m->e[t].bytes;
        for (int at = 0; at <= MAXA; at++)
            for (uint64_t i = 0; i < fl->nrows[at]; i++) {
                cont_t* c = &fl->rows[at][i];
                if (!c->present) continue;
                for (uint32_t j = 0; j < c->n; j++) PUSHV(map_val(f, c->st + (size_t)j * bytes, dtype));
            }
    }
    if (order == 0) {
        /* sequential left fold: reduced = op(f(x), reduced) */
        for (uint64_t i = 0; i < n; i++) acc = op_apply(op, dtype, vals[i], acc);
    } else {
        acc_t tot = tree_reduce(vals, n, op, dtype);
        acc = op_apply(op, dtype, tot, acc);
    }
    free(vals);
    if (is_fdt(dtype)) memcpy(out, &acc.f, 8);
    else memcpy(out, &acc.i, 8);
    return VH_OK;
#undef PUSHV
}
