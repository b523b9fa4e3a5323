/*
 * vahana_hip.h — C ABI of the MI355X (gfx950) execution backend for
 * Vahana.jl's transition step (`apply!`).
 *
 * Every entry point returns an int status (VH_OK == 0, negative on error);
 * the message of the last error on the calling thread is returned by
 * vh_last_error().  Handles are opaque.  Host pointers passed in are
 * borrowed for the duration of the call (the library copies what it keeps);
 * output buffers are owned by the caller and carry an explicit capacity.
 *
 * Reference paths below are relative to the Vahana.jl repository root.
 */
#ifndef VAHANA_HIP_H
#define VAHANA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define VH_OK               0
#define VH_ERR_ASSERT      -1  /* reference: AssertionError (@assert / @mayassert) */
#define VH_ERR_UNREGISTERED -2 /* reference: error() for unregistered types/functions */
#define VH_ERR_OOM         -3  /* device allocation failed */
#define VH_ERR_BUDGET      -4  /* HBM budget exceeded (serving admission) */
#define VH_ERR_DEVICE      -5  /* HIP runtime error */
#define VH_ERR_CAPACITY    -6  /* caller buffer too small / fixed capacity exceeded */
#define VH_ERR_ARG         -7  /* invalid argument */

/* ------------------------------------------------------------ AgentID codec
 * src/Agent.jl:30-118: id = type << 56 | rank << 36 | nr, nr 1-based.
 * This backend runs one rank, so rank == 0 and ids are deterministic. */
#define VH_BITS_TYPE     8
#define VH_BITS_PROCESS  20
#define VH_BITS_AGENTNR  36
#define VH_SHIFT_TYPE    (VH_BITS_PROCESS + VH_BITS_AGENTNR)
#define VH_NR_MASK       ((((uint64_t)1) << VH_BITS_AGENTNR) - 1)

#define VH_MAX_AGENT_TYPES 8
#define VH_MAX_EDGE_TYPES  24
#define VH_MAX_CALL        8

/* ----------------------------------------------------------------- hints
 * Agent hints: src/ModelTypes.jl:94-101.  Edge hints: src/ModelTypes.jl:165-228
 * (NumEdgesOnly and HasEdgeOnly expand to their components, SingleType is
 * implied by a target type, SingleType+SingleEdge needs Stateless+IgnoreFrom). */
#define VH_HINT_IMMORTAL          (1u << 0)
#define VH_HINT_INDEPENDENT       (1u << 1)

#define VH_HINT_STATELESS         (1u << 0)
#define VH_HINT_IGNOREFROM        (1u << 1)
#define VH_HINT_SINGLEEDGE        (1u << 2)
#define VH_HINT_SINGLETYPE        (1u << 3)
#define VH_HINT_IGNORESOURCESTATE (1u << 4)
#define VH_HINT_NUMEDGESONLY      (1u << 5)
#define VH_HINT_HASEDGEONLY       (1u << 6)

/* ----------------------------------------------------- registered transitions
 * The reference's transition function is an arbitrary Julia closure
 * (src/Simulation.jl:720-730).  Here each transition body is a HIP device
 * function compiled into the library and selected by id; role slots bind the
 * body's edge/agent type parameters to the model's registered types. */
enum vh_transition {
    VH_TR_NOOP = 0,          /* return the state unchanged                        */
    VH_TR_SUM_NEIGHBORS = 1, /* foo = sum(n.foo for n in neighborstates_flexible)  */
    VH_TR_DIE_IF = 2,        /* return nothing if predicate(state) else state      */
    VH_TR_ADD_SELF_EDGE = 3, /* add_edge!(sim, id, id, E(params))                  */
    VH_TR_READD_EDGES = 4,   /* add_edges!(sim, id, edges(sim, id, E))              */
    VH_TR_NEIGHBOR_FOO_EDGES = 5, /* test/independent.jl:55-68 body                */
    VH_TR_SPAWN = 6,         /* add_agent!(T(foo + k)) (+ edge newborn -> id)      */
    VH_TR_COUNT_EDGES = 7,   /* foo = num_edges(sim, id, E)                         */
    VH_TR_DIFFUSE = 8,       /* Bool field |= any in-neighbour's (test/raster.jl:38-41) */
    VH_TR_MOVE_TO = 9,       /* value_on_pos: move_to! cell (v,...,v) (test/raster.jl:279-285) */
    VH_TR_REMOVE_EDGES = 10, /* remove_edges! bodies of test/mpi/test_edgetypes.jl:295-445 */
    VH_TR_SIR = 16,          /* SIR contagion (Stateless contact edges)            */
    VH_TR_OPINION_F64 = 17,  /* weighted bounded confidence, Float64 opinions      */
    VH_TR_OPINION_F32 = 18,  /* same, Float32 opinions                              */
    VH_TR_CONTACT = 19,      /* dynamic contact network with deaths/births         */
    VH_TR_COUNT_ = 20,
    VH_TR_USER = 64          /* first id of the bodies from vh_register_body       */
};

/* predicates of VH_TR_DIE_IF (applied to the Int64 field at offset 0) */
enum vh_cmp {
    VH_CMP_ALWAYS = 0, VH_CMP_NEVER = 1, VH_CMP_LT = 2, VH_CMP_GE = 3,
    VH_CMP_EQ = 4, VH_CMP_ODD = 5, VH_CMP_EVEN = 6, VH_CMP_NUM_EDGES_EQ = 7
};

typedef struct vh_test_params {      /* params of the test bodies (ids 0-7) */
    int64_t value;                   /* comparison value / spawn offset     */
    int32_t cmp;                     /* enum vh_cmp                          */
    uint32_t pad;
    uint8_t edge_state[16];          /* edge state bytes for ADD_SELF_EDGE  */
} vh_test_params;

/* VH_TR_REMOVE_EDGES: vh_test_params.value selects the body; the agent's
 * Int64 at offset 0 is its `idx`; role_edge[0] = ET, role_edge[1] = SendId.
 * remove_edges!(sim, to, ET) src/EdgeMethods.jl:527-549 and
 * remove_edges!(sim, from, to, ET) :552-599 act on the write container at
 * the point of the call, in call order with add_edge!. */
enum vh_remove_mode {
    VH_RM_OWN = 0,          /* idx even: remove_edges!(id)                       (:315-317) */
    VH_RM_NEIGHBOR = 1,     /* idx even: remove_edges!(first(neighborids(id)))   (:327-330) */
    VH_RM_READD = 2,        /* f = first(neighborids(id)); remove_edges!(f, id);
                               add_edge!(f, id, edge_state)                       (:340-347) */
    VH_RM_FROM_TWO = 3,     /* n = neighborids(id); remove_edges!(n[1], id), and
                               remove_edges!(n[2], id) unless :SingleEdge         (:387-393) */
    VH_RM_TO_NEIGHBOR = 4,  /* idx even: remove_edges!(id, n) for a neighbour n
                               (the reference draws n at random; the first here)  (:403-408) */
    VH_RM_SEND_ID = 5,      /* add_edge!(id, neighborids(id, ET), SendId)          (:411-414) */
    VH_RM_TO_SENDID = 6,    /* idx < 100: remove_edges!(id, neighborids(id, SendId), ET) (:415-421) */
    VH_RM_ZERO_FROM = 7     /* remove_edges!(AgentID(0), id)                      (:428-430) */
};

typedef struct vh_raster_params {    /* raster bodies (ids 8, 9)                    */
    uint32_t field_offset;           /* DIFFUSE: byte offset of the Bool field       */
    uint32_t cell_type;              /* MOVE_TO: agent type of the raster's cells    */
    uint64_t first_nr;               /* MOVE_TO: agent number of cell (1, ..., 1)    */
    int64_t dims[4];                 /* MOVE_TO: raster size, column-major           */
    uint32_t ndims;                  /* MOVE_TO: 1..4                                 */
    uint32_t flags;                  /* MOVE_TO: 1 edge cell -> agent, 2 agent -> cell */
    uint8_t edge_state[16];          /* MOVE_TO: state of the added edges            */
} vh_raster_params;

typedef struct vh_sir_params {       /* SIR: state {UInt8 health, UInt16 days} */
    double p_inf;                    /* per infected in-neighbour, per step    */
    uint64_t seed;
    uint32_t recovery;               /* days until I -> R                      */
    uint32_t pad;
} vh_sir_params;

typedef struct vh_opinion_params {   /* state Float64 (or Float32), edge Float32 */
    double eps;
} vh_opinion_params;

typedef struct vh_contact_params {   /* state {UInt8 health, Float32 susceptibility} */
    double beta;                     /* infection prob. scale per infected contact  */
    double gamma;                    /* recovery probability per step               */
    double p_death;                  /* death (and 1:1 birth) probability per step  */
    uint64_t seed;
    uint32_t n_contacts;             /* partners drawn per agent and step           */
    uint32_t max_redraw;             /* redraws per partner when the draw is dead   */
} vh_contact_params;

/* model-state enums shared by SIR and contact */
#define VH_S 0
#define VH_I 1
#define VH_R 2

/* ------------------------------------------------------------- apply! desc
 * apply!(sim, func, call, read, write; add_existing, with_edge)
 * src/Simulation.jl:720-821.  Type lists are bit masks over registered type
 * ids (agent type ids are 1-based like src/ModelTypes.jl:86; bit i <=> id i),
 * except `call`, whose order fixes the side-effect order. */
/* A field of the body's parameter blob taken from a simulation value (see
 * vh_sim_set_value) when apply! is enqueued: param(sim, name) and
 * get_global(sim, name) read inside a transition (src/Simulation.jl:587,
 * src/Global.jl:10), both constant while a transition runs (set_param! only
 * before finish_init!, set_global! never inside a transition). The value is
 * converted to the field's type as Julia's convert does (an integer field
 * takes only integral values in its range). */
typedef struct vh_value_bind {
    uint32_t offset;                 /* byte offset of the field in the blob   */
    uint8_t  kind;                   /* VH_VALUE_PARAM or VH_VALUE_GLOBAL      */
    uint8_t  dtype;                  /* enum vh_dtype of the field             */
    uint16_t slot;
} vh_value_bind;

typedef struct vh_apply_desc {
    uint32_t transition;             /* enum vh_transition                    */
    uint32_t n_call;
    uint8_t  call[VH_MAX_CALL];      /* agent type ids in call order           */
    uint32_t read_agents, read_edges;    /* masks                              */
    uint32_t write_agents, write_edges;
    uint32_t add_existing_agents, add_existing_edges;
    int32_t  with_edge;              /* edge type id or -1                     */
    int8_t   role_agent[4];          /* body's agent-type parameters (-1 = unused) */
    int8_t   role_edge[4];           /* body's edge-type parameters  (-1 = unused) */
    const void* params;              /* body parameter blob (copied)           */
    uint32_t params_bytes;
    uint32_t n_bind;                 /* blob fields taken from simulation values */
    const vh_value_bind* bind;
} vh_apply_desc;

/* ------------------------------------------------------------- mapreduce
 * src/AgentMethods.jl:533-565 (agents), src/EdgeMethods.jl:972-1001 (edges),
 * init values of val4empty src/Helpers.jl:44-81. */
enum vh_op { VH_OP_ADD = 0, VH_OP_MUL = 1, VH_OP_MAX = 2, VH_OP_MIN = 3,
             VH_OP_AND = 4, VH_OP_OR = 5 };
enum vh_dtype { VH_DT_I64 = 0, VH_DT_F64 = 1, VH_DT_BOOL = 2, VH_DT_F32 = 3,
                VH_DT_U8 = 4, VH_DT_U16 = 5, VH_DT_U32 = 6, VH_DT_I32 = 7,
                VH_DT_U64 = 8 };
enum vh_map_kind { VH_MAP_FIELD = 0,      /* f(x) = x.field                     */
                   VH_MAP_FIELD_EQ = 1,   /* f(x) = x.field == value (Bool/Int) */
                   VH_MAP_ONE = 2 };      /* f(x) = 1                           */
/* A computed f for mapreduce (the reference takes any function of the
 * state): a postfix program over one record. Values are Int64, Float64 or
 * Bool; FIELD pushes the field at byte offset `arg` of type `dtype` (floats
 * as Float64, Bool as Bool, other integers as Int64); CONST_I / CONST_F /
 * CONST_B push `arg` (Int64 bits, Float64 bits, 0 or 1). ADD SUB MUL MIN
 * MAX: Float64 if an operand is a float, else Int64 (wrapping); DIV: Float64
 * (Julia's /); MIN / MAX on floats as Julia's min / max (NaN wins, -0.0 <
 * +0.0); NEG / ABS keep the type (Bool as Int64); comparisons compare as
 * Float64 if an operand is a float, else as Int64, and give Bool; AND / OR /
 * NOT are logical (nonzero is true) and give Bool; SELECT pops c, a, b
 * (pushed in that order) and gives c ? a : b (Float64 if a or b is a
 * float). The program's value is converted to the reduction's dtype: a
 * float to an integer dtype truncates (NaN gives 0, values out of range
 * saturate), any value to Bool is != 0. */
enum vh_expr_code { VH_EX_FIELD = 0, VH_EX_CONST_I = 1, VH_EX_CONST_F = 2, VH_EX_CONST_B = 3,
                    VH_EX_ADD = 10, VH_EX_SUB = 11, VH_EX_MUL = 12, VH_EX_DIV = 13, VH_EX_MIN = 14,
                    VH_EX_MAX = 15, VH_EX_NEG = 20, VH_EX_ABS = 21, VH_EX_NOT = 22, VH_EX_EQ = 30,
                    VH_EX_NE = 31, VH_EX_LT = 32, VH_EX_LE = 33, VH_EX_GT = 34, VH_EX_GE = 35,
                    VH_EX_AND = 40, VH_EX_OR = 41, VH_EX_SELECT = 50 };
#define VH_EXPR_MAX_OPS   64
#define VH_EXPR_MAX_DEPTH 16
typedef struct vh_expr_op {
    uint32_t code;                   /* enum vh_expr_code                     */
    uint32_t dtype;                  /* FIELD: enum vh_dtype of the field     */
    uint64_t arg;                    /* FIELD: byte offset; CONST_*: value bits */
} vh_expr_op;

typedef struct vh_map_desc {
    uint32_t kind;                   /* enum vh_map_kind                      */
    uint32_t offset;                 /* byte offset of the field in the state */
    uint32_t field_dtype;            /* enum vh_dtype of the field            */
    uint32_t pad;
    int64_t  value;                  /* comparison value for FIELD_EQ         */
} vh_map_desc;

typedef struct vh_model* vh_model_t;
typedef struct vh_sim*   vh_sim_t;

/* ------------------------------------------------------------- model */
/* ModelTypes() src/ModelTypes.jl:1-74 */
int vh_model_create(vh_model_t* out);
int vh_model_destroy(vh_model_t m);
/* register_agenttype! src/ModelTypes.jl:81-108 */
int vh_register_agenttype(vh_model_t m, uint32_t hints, uint32_t state_bytes,
                          uint8_t* type_id);
/* register_edgetype! src/ModelTypes.jl:158-230 (target_type < 0: none) */
int vh_register_edgetype(vh_model_t m, uint32_t hints, uint32_t state_bytes,
                         int32_t target_type, int64_t size, uint8_t* edge_type_id);
/* resolved hints after normalisation (src/ModelTypes.jl:181-205) */
int vh_edgetype_hints(vh_model_t m, uint8_t edge_type_id, uint32_t* hints);

/* ------------------------------------------------------------- simulation */
/* create_simulation src/Simulation.jl:261-339 (allocates nothing on the
 * device until finish_init!) */
int vh_sim_create(vh_model_t m, vh_sim_t* out);
/* finish_simulation! src/Simulation.jl:551-585 */
int vh_sim_destroy(vh_sim_t s);
/* copy_simulation src/Simulation.jl:500-510 (device deep copy) */
int vh_sim_copy(vh_sim_t s, vh_sim_t* out);
/* add_agent! / add_agents! during initialisation
 * src/AgentMethods.jl:65-89, src/Agent.jl:174-180 */
int vh_add_agents(vh_sim_t s, uint8_t type_id, const void* states, uint64_t n,
                  uint64_t* ids_out);
/* add_edge! / add_edges! during initialisation, in call order
 * src/EdgeMethods.jl:388-523, src/Edge.jl:81-97 (states may be NULL for
 * stateless types) */
int vh_add_edges(vh_sim_t s, uint8_t edge_type_id, const uint64_t* from,
                 const uint64_t* to, const void* states, uint64_t n);

/* Bulk init-phase edges (no single reference counterpart: the edges that
 * add_edge!, src/EdgeMethods.jl:388-523, would add one by one for a directed
 * G(n, p) graph, target by target). Edge type `et` of a model with one agent
 * type gets, at finish_init!, a Binomial(n - 1, p) number of in-edges per
 * agent t (p = mean_deg / (n - 1), n = agents added) from distinct uniform
 * sources != t, in draw order; a Float32 edge state is the weight
 * (2 (h >> 41) + 1) 2^-24, h = vh_hash4(weight_seed, k, 0, 6 << 40) for
 * the k-th edge (0-based, target order). vh_graph_gnp and vh_fill_open01f
 * give the same edges on the host. No other edges may be added to `et`. */
int vh_add_edges_gnp(vh_sim_t sim, uint8_t et, double mean_deg, uint64_t graph_seed,
                     uint64_t weight_seed);
/* Bulk init-phase edges for add_graph! of a generated Barabasi-Albert graph
 * (src/GraphsSupport.jl:34-57; the edges add_edge!, src/EdgeMethods.jl:
 * 388-523, would add): the undirected pairs vh_graph_ba(n, m, graph_seed)
 * gives over the n agents of the edge type's target type, each added in
 * both directions (pair k: edge 2k a_k -> b_k, edge 2k + 1 b_k -> a_k); a
 * Float32 edge state is the weight of vh_fill_open01f(weight_seed, 0,
 * 6 << 40) at that edge position. Generated on the device at finish_init!
 * (and numbered with the other init edges by vh_finish_init_relabel). No
 * other edges may be added to `et`; 1 <= m <= 255, 2 m n < 2^32. */
int vh_add_edges_ba(vh_sim_t sim, uint8_t et, uint32_t m, uint64_t graph_seed, uint64_t weight_seed);

/* connect_raster_neighbors! src/Raster.jl:134-164 for a raster whose
 * prod(dims) cells are the agents first_nr, first_nr + 1, ... of cell_type
 * in column-major order (as add_raster! adds them): for every cell `org`
 * and stencil offset `s` an edge org -> org + s (wrapped per dimension when
 * periodic, dropped outside the raster otherwise), all with `edge_state`.
 * Stencil as src/Raster.jl:80-96 (metric 0 Chebyshev, 1 Euclidean, 2
 * Manhattan; at most 128 offsets). Generated on the device at finish_init!;
 * the CSR equals adding the edges one by one, cell by cell. No other edges
 * may be added to `et`. */
int vh_add_edges_stencil(vh_sim_t sim, uint8_t et, uint32_t cell_type, uint64_t first_nr, uint32_t ndims,
                         const int64_t* dims, double distance, int metric, int periodic,
                         const uint8_t* edge_state);
/* A transition body loaded at run time (the reference's apply! takes any
 * function, src/Simulation.jl:720-730): kernel `kernel` of the code object
 * file `code_object`, defined with VH_BODY_KERNEL of include/vahana_device.h
 * and compiled for gfx950 (hipcc --genco). `flags`: VH_BODY_ADDS_EDGES when
 * the body calls add_edge! / remove_edges! (entries of edge role 0),
 * VH_BODY_ADDS_AGENTS when it calls add_agent! (agent role 0); such bodies
 * run twice per apply! (count, then write). *transition_id receives the id
 * to put in vh_apply_desc.transition (VH_TR_USER + k); registering the same
 * kernel again gives the same id. At most VH_MAX_USER_BODIES bodies. */
#define VH_BODY_ADDS_EDGES  1u
#define VH_BODY_ADDS_AGENTS 2u
#define VH_MAX_USER_BODIES  64
int vh_register_body(const char* code_object, const char* kernel, uint32_t flags, uint32_t* transition_id);
/* finish_init! src/Simulation.jl:403-476: upload and build the CSR */
int vh_finish_init(vh_sim_t s);
/* finish_init! with the agents of type t renumbered in breadth-first order
 * of the init edges among them (the reference renumbers agents when it
 * distributes them, src/Simulation.jl:414-453, and returns the mapping with
 * `return_idmapping`): new_nr[k] receives the new number of the agent added
 * as number k + 1. States travel with their agents and every target keeps
 * its in-edges in insertion order, so results map one to one. AgentIDs
 * stored inside agent or edge states are not renumbered. */
int vh_finish_init_relabel(vh_sim_t sim, uint8_t t, uint64_t* new_nr);
/* apply! src/Simulation.jl:720-821, enqueued on the simulation's stream */
int vh_apply(vh_sim_t s, const vh_apply_desc* desc);
/* apply! repeated `steps` times with the same desc (one host call) */
int vh_apply_n(vh_sim_t s, const vh_apply_desc* desc, uint32_t steps);
/* apply! of one transition to many simulations (no reference counterpart:
 * the reference runs one simulation per process): simulation i takes part in
 * the first steps[i] steps, and the one-pass opinion bodies of all
 * simulations of a step run as one launch per state type (simulations whose
 * state needs the two-pass gather, and other bodies, run their own kernels).
 * All simulations must run on one stream (vh_set_stream) and have the type
 * ids the desc names. Asynchronous like vh_apply; a call first waits for the
 * batch's previous call. */
typedef struct vh_batch* vh_batch_t;
int vh_batch_create(vh_batch_t* out);
int vh_batch_destroy(vh_batch_t b);
int vh_batch_apply(vh_batch_t b, const vh_sim_t* sims, uint32_t n, const vh_apply_desc* desc,
                   const uint32_t* steps);
/* wait for the simulation's stream */
int vh_sync(vh_sim_t s);
/* build, on the simulation's stream and waiting for it, what the first
 * apply! of the opinion transition (VH_TR_OPINION_F64, or _F32 with f32 != 0)
 * of agent type t over edge type et would build: its degree bins and, for a
 * state that takes the two-pass gather, its gather plan (no reference
 * counterpart: for servers that build simulations on one stream and step
 * them on another, so that the plan's host round trips stay off the step
 * stream). Kept until the edges change; no effect on results. */
int vh_opinion_prepare(vh_sim_t s, uint8_t t, uint8_t et, int f32);
/* attach an external hipStream_t (NULL: the library's own stream) */
int vh_set_stream(vh_sim_t s, void* hip_stream);
/* a non-blocking hipStream_t at default priority for vh_set_stream,
 * vh_set_io_stream and vh_ensemble_set_streams (no reference counterpart) */
int vh_stream_create(void** out);
int vh_stream_destroy(void* hip_stream);
/* stream for vh_read_agents (no reference counterpart): the copies wait for
 * the simulation's last apply! only, so a caller that runs many simulations
 * on one shared compute stream reads finished ones while the others run;
 * NULL: reads use the compute stream */
int vh_set_io_stream(vh_sim_t s, void* hip_stream);
/* param / set_param! (src/Simulation.jl:587-630, register_param!
 * src/ModelTypes.jl:244-266) and get_global / set_global! / modify_global!
 * (src/Global.jl:10-59, register_global! src/ModelTypes.jl:268-288) as
 * numbered scalar slots per simulation; the host layer keeps the names and
 * non-scalar globals (the time series of push_global!). `value` points to one
 * value of `dtype` (enum vh_dtype). A parameter can be set only before
 * finish_init!, as set_param! asserts. */
#define VH_MAX_VALUES   32
#define VH_VALUE_PARAM  0
#define VH_VALUE_GLOBAL 1
int vh_sim_set_value(vh_sim_t s, int kind, uint32_t slot, int dtype, const void* value);
/* *dtype receives the stored type, `value` (8 bytes) the value; VH_ERR_ARG
 * for a slot never set */
int vh_sim_get_value(vh_sim_t s, int kind, uint32_t slot, int* dtype, void* value);
/* sim.num_transitions (src/Simulation.jl:473,816) */
int vh_num_transitions(vh_sim_t s, uint64_t* out);
/* read_snapshot! src/HDF5.jl:1594-1646 (_read_agents_restore! :801-850), after
 * finish_init!: the died flags of the n slots of agent type t; the reuse
 * stack becomes the died slots in ascending order, as the reference
 * rebuilds it */
int vh_restore_died(vh_sim_t sim, uint8_t t, const uint8_t* died, uint64_t n);
/* read_snapshot! sets sim.num_transitions to the snapshot's transition */
int vh_set_num_transitions(vh_sim_t sim, uint64_t n);
/* last_change of an agent (kind 0) or edge (kind 1) type: the transition that
 * last wrote it (finish_write! src/AgentMethods.jl:433, src/EdgeMethods.jl:682)
 * or whose deaths removed some of its edges (src/AgentMethods.jl:347-353); 0
 * after finish_init!. write_snapshot skips types unchanged since their last
 * write (src/HDF5.jl:497-503,592-598); read_snapshot! restores the value. */
int vh_last_change(vh_sim_t sim, int kind, uint8_t type_id, uint64_t* out);
int vh_set_last_change(vh_sim_t sim, int kind, uint8_t type_id, uint64_t value);
/* mapreduce(sim, f, op, T; init, datatype) — kind 0: agents, 1: edges */
int vh_mapreduce(vh_sim_t s, int kind, uint8_t type_id, const vh_map_desc* f,
                 int op, int dtype, const void* init, void* out);
/* mapreduce(sim, f, op, T; init, datatype) with f a vh_expr_op program of
 * nops operations (at most VH_EXPR_MAX_OPS, stack depth at most
 * VH_EXPR_MAX_DEPTH, fields inside the record); same reduction and results
 * as vh_mapreduce otherwise. VH_ERR_ARG for an invalid program. */
int vh_mapreduce_expr(vh_sim_t s, int kind, uint8_t type_id, const vh_expr_op* prog, uint32_t nops,
                      int op, int dtype, const void* init, void* out);
/* vh_mapreduce's reduction without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (tight
 * serving loops, code under a HIP graph capture, harnesses that forbid
 * hipMalloc). Same arguments and results as vh_mapreduce; the block partials
 * live in a fixed library array of (2^17 + 256) 8-byte words per GPU, which
 * limits a call to 2^28 records (VH_ERR_CAPACITY beyond). Calls that overlap
 * on one GPU share that array and race. */
int vh_mapreduce_noalloc(vh_sim_t s, int kind, uint8_t type_id, const vh_map_desc* f,
                         int op, int dtype, const void* init, void* out);
/* num_agents src/Agent.jl:324-343 */
int vh_num_agents(vh_sim_t s, uint8_t type_id, uint64_t* out);
/* num_edges(sim, T) src/Edge.jl:373-389 */
int vh_num_edges(vh_sim_t s, uint8_t edge_type_id, uint64_t* out);
/* raw slot view of read state + died flags (all_agents src/Agent.jl:234-260
 * filters it); *nslots receives length(readstate) */
int vh_read_agents(vh_sim_t s, uint8_t type_id, void* states, uint8_t* died,
                   uint64_t cap, uint64_t* nslots);
/* every edge as (to, from, state) in container order (all_edges
 * src/EdgeMethods.jl:1004-1027; order of src/EdgeIterator.jl:45-106 with
 * ascending (type, nr) target keys); from/states may be NULL */
int vh_read_edges(vh_sim_t s, uint8_t edge_type_id, uint64_t* to,
                  uint64_t* from, void* states, uint64_t cap, uint64_t* n);
/* edges / neighborids / edgestates / num_edges of one target outside a
 * transition (src/EdgeMethods.jl:704-892) */
int vh_edges_of(vh_sim_t s, uint8_t edge_type_id, uint64_t to, uint64_t* from,
                void* states, uint64_t cap, uint64_t* n);
/* agentstate outside a transition src/AgentMethods.jl:91-154 */
int vh_agentstate(vh_sim_t s, uint64_t id, uint8_t type_id, void* out);
/* bytes of device memory held by the simulation */
int vh_device_bytes(vh_sim_t s, uint64_t* out);
/* last error message on this thread */
const char* vh_last_error(void);
/* ------------------------------------------------------------- options
 * Process-wide tuning options of the transition kernels (no reference
 * counterpart: the reference has no device kernels). Every apply! and
 * ensemble step reads the values set last; set them between calls, not while
 * another thread applies. Results do not depend on them (every setting gives
 * bit-identical states), only the kernels chosen do. */
typedef struct vh_options {
    int64_t opinion_plan_bytes;   /* opinion states gathered in two passes through a plan from a
                                     state vector of at least this many bytes (< 0: never);
                                     default 16 MiB */
    int64_t opinion_part_bytes;   /* source window of the plan in bytes (0 or more than 160 KiB:
                                     160 KiB, one CU's LDS; widened to make at most 1024
                                     windows) */
    int32_t sir_expose;           /* SIR exposure pass: 1 on (default), 0 off */
    uint32_t pad;
    int64_t sir_expose_min;       /* agents from which a SIR simulation or ensemble uses the
                                     exposure pass; default 65536 */
} vh_options;
int vh_get_options(vh_options* out);
int vh_set_options(const vh_options* in);

/* enable_asserts src/Vahana.jl:42-48: with asserts on (the default) every
 * apply! waits for its kernels and reports device-side assertion flags
 * (permissions, removed agents, capacity) as VH_ERR_ASSERT / VH_ERR_CAPACITY;
 * with asserts off apply! stays asynchronous unless it has births or deaths */
int vh_enable_asserts(int enable);
/* test hooks (no reference counterpart): the next n device allocations of
 * the library fail as out of memory (an injected allocation failure); OR
 * `flags` into a simulation's device error word, as a failing body would
 * (an injected device error: the next checked apply! reports it) */
int vh_debug_fail_allocs(uint32_t n);
int vh_debug_set_device_error(vh_sim_t sim, uint32_t flags);

/* ------------------------------------------------------ ensembles
 * Many independent SIR simulations advanced by one segmented launch per
 * step: parameter sweeps (every member bit-identical to the same simulation
 * run alone through apply!) and the step batches of the serving loop.
 * Members own their device buffers and may join or leave between steps.
 * No counterpart in the reference (it runs one simulation per MPI job). */
typedef struct vh_ensemble* vh_ensemble_t;
typedef struct vh_sir_member {
    uint64_t n;            /* agents                                          */
    double mean_deg;       /* G(n, p) with p = mean_deg / (n - 1)            */
    double p_inf;          /* infection probability per infected in-neighbour */
    uint64_t seed;         /* transition and initial-state seed              */
    uint64_t graph_seed;   /* vh_graph_gnp seed                              */
    uint32_t recovery;     /* days until I -> R                              */
    uint32_t n_init;       /* expected initial I: each agent with n_init / n */
} vh_sir_member;
int vh_ensemble_create(vh_ensemble_t* out);
int vh_ensemble_destroy(vh_ensemble_t e);
/* builds the member's graph and state on a second stream of the ensemble,
 * so it may run while launched steps of the other members are in flight */
int vh_ensemble_add_sir(vh_ensemble_t e, const vh_sir_member* m, uint32_t* member_id);
/* waits for the member's last launched step only */
int vh_ensemble_remove(vh_ensemble_t e, uint32_t member_id);
/* the member applies steps up to last_step and then stays (0: no limit) */
int vh_ensemble_set_horizon(vh_ensemble_t e, uint32_t member_id, uint64_t last_step);
/* block size (128..1024) and XCD-aware block order of the step launch */
int vh_ensemble_set_config(vh_ensemble_t e, uint32_t block, uint32_t xcd_remap);
/* advance every member; use_graph: capture the steps into one hipGraph.
 * Asynchronous: returns once the launches are enqueued (the call first waits
 * for this ensemble's previous launches, not for other work on its stream) */
int vh_ensemble_step(vh_ensemble_t e, uint32_t steps, int use_graph);
int vh_ensemble_sync(vh_ensemble_t e);
/* run the step launches on `step_stream` and member builds and reads on
 * `aux_stream` (both caller-owned hipStream_t; both NULL: the ensemble's own) */
int vh_ensemble_set_streams(vh_ensemble_t e, void* step_stream, void* aux_stream);
/* member state (4-byte SIR records) and its next step number; waits for the
 * member's last launched step only (a member past its horizon is read while
 * later steps of the others run) */
int vh_ensemble_read(vh_ensemble_t e, uint32_t member_id, void* states, uint64_t cap, uint64_t* n,
                     uint64_t* step);
int vh_ensemble_info(vh_ensemble_t e, uint64_t* members, uint64_t* agents, uint64_t* edges,
                     uint64_t* bytes);
/* agents per thread (1, 2 or 4) of the members of each size class
 * ([0, 10^4), [10^4, 10^5), [10^5, ...) agents): a member's logical blocks
 * cover block * apt agents */
int vh_ensemble_set_tiles(vh_ensemble_t e, const uint32_t* apt3);
/* times, on a snapshot restored afterwards, every (block, order) candidate
 * with one agent per thread, then 2 and 4 agents per thread for each size
 * class present; keeps the fastest (class_apt: 3 entries, may be NULL) */
int vh_ensemble_autotune(vh_ensemble_t e, uint32_t trial_steps, uint32_t* block, uint32_t* xcd_remap,
                         uint32_t* class_apt, float* ms_per_step);

/* ------------------------------------------------------ graph generators
 * Host-side sources for add_graph! (src/GraphsSupport.jl:34-57) with a
 * counter-hash stream so both the device path and any checker receive the
 * same edge list.  Pairs are 1-based agent numbers.
 * Directed G(n, p): for each target ascending, sources ascending. */
int vh_graph_er(uint64_t n, double p, uint64_t seed, uint32_t* from,
                uint32_t* to, uint64_t cap, uint64_t* m);
/* Directed G(n, p), p = mean_deg / (n - 1): target t draws its in-degree
 * from Binomial(n - 1, p) by an inverse-CDF table, then that many distinct
 * uniform sources; edges per target in draw order, targets ascending.  The
 * ensemble builds the identical graph on the device. */
int vh_graph_gnp(uint64_t n, double mean_deg, uint64_t seed, uint32_t* from, uint32_t* to,
                 uint64_t cap, uint64_t* m);
/* Barabasi-Albert, m0 = m clique, undirected pairs in creation order
 * (count = m(m-1)/2 + (n-m)m) */
int vh_graph_ba(uint64_t n, uint32_t m, uint64_t seed, uint32_t* a, uint32_t* b,
                uint64_t cap, uint64_t* count);

/* Float32 weights in (0, 1): out[i] = (2 (h >> 41) + 1) 2^-24 with
 * h = vh_hash4(seed, i, b, c) */
int vh_fill_open01f(uint64_t seed, uint64_t b, uint64_t c, uint64_t n, float* out);

/* Float64 in [0, 1) per agent: out[i] = (h >> 11) 2^-53 with
 * h = vh_hash4(seed, AgentID(type, i + 1), b, c) (the opinion and other
 * per-agent initial values) */
int vh_fill_agent_u01(uint64_t seed, uint32_t type, uint64_t b, uint64_t c, uint64_t n, double* out);

/* breadth-first numbering of vertices 1..n over m undirected edges
 * (a[k], b[k]): adjacency in edge order, the first root the vertex of
 * highest degree (lowest number among equals), further components from their
 * lowest-numbered vertex; new_nr[v - 1] = 1-based position of v (the order
 * vh_finish_init_relabel applies) */
int vh_bfs_order(uint64_t n, uint64_t m, const uint32_t* a, const uint32_t* b, uint64_t* new_nr);

/* ------------------------------------------------------ counter-hash RNG
 * SplitMix64 finaliser chained over (seed, a, b, c); the same function is
 * compiled into the device bodies. */
uint64_t vh_hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c);

#ifdef __cplusplus
}
#endif
#endif /* VAHANA_HIP_H */
