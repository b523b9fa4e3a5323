/*
 * vh_oracle.h — CPU restatement of Vahana.jl's transition step, used ONLY
 * as a checker by the tests, by __graft_entry__.smoke() and as the CPU
 * baseline leg of bench.py.  It is test infrastructure: the product library
 * (libvahana_hip.so) never links, loads or calls it.
 *
 * Entry points mirror include/vahana_hip.h with an `or_` prefix and share its
 * constants and parameter structs (the interface, not the implementation).
 */
#ifndef VH_ORACLE_H
#define VH_ORACLE_H
#include "../include/vahana_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_model or_model;
typedef struct or_sim or_sim;

int or_model_create(or_model** out);
int or_model_destroy(or_model* m);
int or_register_agenttype(or_model* m, uint32_t hints, uint32_t bytes, uint8_t* id);
int or_register_edgetype(or_model* m, uint32_t hints, uint32_t bytes,
                         int32_t target, int64_t size, uint8_t* id);
int or_edgetype_hints(or_model* m, uint8_t et, uint32_t* hints);
int or_sim_create(or_model* m, or_sim** out);
int or_sim_destroy(or_sim* s);
int or_sim_copy(or_sim* s, or_sim** out);
int or_add_agents(or_sim* s, uint8_t t, const void* states, uint64_t n, uint64_t* ids);
int or_add_edges(or_sim* s, uint8_t et, const uint64_t* from, const uint64_t* to,
                 const void* states, uint64_t n);
int or_finish_init(or_sim* s);
int or_apply(or_sim* s, const vh_apply_desc* d);
int or_sim_set_value(or_sim* s, int kind, uint32_t slot, int dtype, const void* value);
int or_sim_get_value(or_sim* s, int kind, uint32_t slot, int* dtype, void* value);
int or_num_transitions(or_sim* s, uint64_t* out);
int or_last_change(or_sim* s, int kind, uint8_t t, uint64_t* out);
int or_set_last_change(or_sim* s, int kind, uint8_t t, uint64_t v);
int or_restore_died(or_sim* s, uint8_t t, const uint8_t* died, uint64_t n);
int or_set_num_transitions(or_sim* s, uint64_t n);
/* order: 0 = the reference's sequential left fold, 1 = the device's fixed tree */
/* bodies supplied by the caller (the checker's counterpart of
 * vh_register_body): fn returns 1 with the new state in `out` (a copy of
 * `st` on entry) or 0 for `nothing`; it reaches the simulation through the
 * or_ctx_* accessors (the reference's transition API) */
typedef struct or_ctx or_ctx;
typedef int (*or_body_fn)(or_ctx* c, const uint8_t* st, uint64_t id, uint8_t* out);
int or_register_body(or_body_fn fn, uint32_t flags, uint32_t* id);
const void* or_ctx_params(or_ctx* c);
int or_ctx_role_edge(or_ctx* c, int k);
int or_ctx_role_agent(or_ctx* c, int k);
/* edges(sim, id, ET): count, sources and states in insertion order */
uint32_t or_ctx_in_edges(or_ctx* c, int et, uint64_t id, const uint64_t** from, const uint8_t** states);
const uint8_t* or_ctx_agentstate(or_ctx* c, uint64_t id);
void or_ctx_add_edge(or_ctx* c, int et, uint64_t from, uint64_t to, const void* st);
uint64_t or_ctx_add_agent(or_ctx* c, int t, const void* st);
void or_ctx_remove_edges(or_ctx* c, int et, uint64_t from, uint64_t to, int all);
int or_ctx_failed(or_ctx* c);
int or_mapreduce_expr(or_sim* s, int kind, uint8_t t, const vh_expr_op* prog, uint32_t nops, int op,
                      int dtype, const void* init, void* out, int order);
int or_mapreduce(or_sim* s, int kind, uint8_t t, const vh_map_desc* f, int op,
                 int dtype, const void* init, void* out, int order);
int or_num_agents(or_sim* s, uint8_t t, uint64_t* out);
int or_num_edges(or_sim* s, uint8_t et, uint64_t* out);
int or_read_agents(or_sim* s, uint8_t t, void* states, uint8_t* died, uint64_t cap,
                   uint64_t* nslots);
int or_read_edges(or_sim* s, uint8_t et, uint64_t* to, uint64_t* from, void* states,
                  uint64_t cap, uint64_t* n);
int or_edges_of(or_sim* s, uint8_t et, uint64_t to, uint64_t* from, void* states,
                uint64_t cap, uint64_t* n);
int or_agentstate(or_sim* s, uint64_t id, uint8_t t, void* out);
/* reuse stack (readreuseable) of a mortal type, bottom to top */
int or_reuse_stack(or_sim* s, uint8_t t, uint64_t* out, uint64_t cap, uint64_t* n);
const char* or_last_error(void);
uint64_t or_hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c);
/* canonical-order weighted sum of the opinion body, exposed for the tests */
void or_canon_sum_f64(const double* num_terms, const double* den_terms, uint32_t d,
                      double* num, double* den);

#ifdef __cplusplus
}
#endif
#endif
