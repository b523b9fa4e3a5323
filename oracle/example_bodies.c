/* example_bodies.c — the checker's restatements of the run-time bodies in
 * vahana.jl_amd/examples/bodies.hip, over the or_ctx accessors. Test
 * infrastructure like the rest of oracle/: tests register them with
 * or_register_body and compare the device's results with them. */
#include <math.h>
#include <string.h>

#include "vh_oracle.h"

/* state.field += delta for a Float64 field (test/hdf5_common.jl:107-109) */
typedef struct { uint32_t offset, pad; double delta; } ex_add_f64_params;

int or_ex_add_f64(or_ctx* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)st; (void)id;
    const ex_add_f64_params* p = (const ex_add_f64_params*)or_ctx_params(c);
    double v;
    memcpy(&v, out + p->offset, 8);
    v = v + p->delta;
    memcpy(out + p->offset, &v, 8);
    return 1;
}

/* remove_some_rasternodes, test/hdf5_common.jl:129-134 */
typedef struct { uint32_t f_offset, pad; double modulus, threshold; } ex_raster_params;

/* Julia's mod1 for floats: the remainder with the divisor's sign, the
 * divisor where that is zero */
static double jl_mod1(double x, double m) {
    double r = fmod(x, m);
    if (r != 0.0 && signbit(r) != signbit(m)) r = r + m;
    return r == 0.0 ? m : r;
}

int or_ex_remove_some_rasternodes(or_ctx* c, const uint8_t* st, uint64_t id, uint8_t* out) {
    (void)st; (void)out;
    const ex_raster_params* p = (const ex_raster_params*)or_ctx_params(c);
    const int et = or_ctx_role_edge(c, 0);
    const uint64_t* from;
    const uint8_t* states;
    uint32_t n = or_ctx_in_edges(c, et, id, &from, &states);
    if (!n) return 1;
    const uint64_t f0 = from[0];
    int64_t a;
    memcpy(&a, states, 8);
    const uint8_t* ns = or_ctx_agentstate(c, f0);
    if (!ns) return 1;
    const int64_t a2 = a * 2;
    or_ctx_add_edge(c, et, f0, id, &a2);
    double f;
    memcpy(&f, ns + p->f_offset, 8);
    return !(jl_mod1(f, p->modulus) > p->threshold);
}
