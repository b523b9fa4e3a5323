// bodies.hip — transition bodies written against include/vahana_device.h and
// loaded at run time (vh_register_body), as a model author would add them:
//
//   hipcc --genco --offload-arch=gfx950 -O3 -ffp-contract=off -I include \
//         vahana.jl_amd/examples/bodies.hip -o bodies.co
//
// then vahana_amd.register_body("bodies.co", "vh_ex_add_f64") etc.
#include "vahana_device.h"

// state.field += delta for a Float64 field (the Agent(state.inner, state.f -
// 2) transition of test/hdf5_common.jl:107-109 with delta = -2)
struct AddF64Params {
    uint32_t offset;   // byte offset of the field
    uint32_t pad;
    double delta;
};
struct AddF64 {
    __device__ static bool run(const DCtx* c, const uint8_t*, uint64_t, Emitter&, uint8_t* out) {
        const AddF64Params* p = (const AddF64Params*)c->params;
        double v;
        memcpy(&v, out + p->offset, 8);
        v = v + p->delta;
        memcpy(out + p->offset, &v, 8);
        return true;
    }
};
VH_BODY_KERNEL(vh_ex_add_f64, AddF64)

// remove_some_rasternodes, test/hdf5_common.jl:129-134:
//     nstate = first(neighborstates(sim, id, RasterEdge, Agent))
//     e = first(edges(sim, id, RasterEdge))
//     add_edge!(sim, e.from, id, RasterEdge(e.state.a * 2))
//     mod1(nstate.f, 8) > 4.5 ? nothing : state
// Edge role 0: RasterEdge (its state an Int64 at offset 0); the source's
// Float64 at `f_offset`.
struct RasterParams {
    uint32_t f_offset;
    uint32_t pad;
    double modulus;     // 8
    double threshold;   // 4.5
};
// Julia's mod1(x, m) for floats: mod(x, m) (the remainder with the sign of
// m), m where that is zero
__device__ __forceinline__ double jl_mod1(double x, double m) {
    double r = fmod(x, m);
    if (r != 0.0 && signbit(r) != signbit(m)) r = r + m;
    return r == 0.0 ? m : r;
}
struct RemoveSomeRasterNodes {
    __device__ static bool run(const DCtx* c, const uint8_t*, uint64_t id, Emitter& em, uint8_t*) {
        const RasterParams* p = (const RasterParams*)c->params;
        const int et = c->role_e[0];
        uint32_t b, e;
        if (!d_row(c, et, id, b, e)) return true;   // `first` of no edges: nothing to do here
        const DEdge& E = c->e[et];
        const uint64_t from = vh_src_id(E, b);
        const uint8_t* ns = d_state(c, from);
        if (!ns) return true;
        int64_t a;
        memcpy(&a, E.st + (uint64_t)b * E.bytes, 8);
        const int64_t a2 = a * 2;
        em.edge(from, id, &a2);
        double f;
        memcpy(&f, ns + p->f_offset, 8);
        return !(jl_mod1(f, p->modulus) > p->threshold);
    }
};
VH_BODY_KERNEL(vh_ex_remove_some_rasternodes, RemoveSomeRasterNodes)
