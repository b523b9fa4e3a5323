// vh_userbody.hip — transition bodies loaded at run time: a body written
// against include/vahana_device.h and compiled into a code object is
// registered with vh_register_body and then applied like a built-in body
// (the reference applies any Julia function, src/Simulation.jl:720-730).
// The library loads the code object on each device the first time a
// simulation there applies the body, and launches its kernel with the
// arguments of the built-in agent loop.
#include <mutex>
#include <string>
#include <vector>

#include "vh_engine.h"

using namespace vh;

namespace {

constexpr int UB_MAX_DEVICES = 64;

struct UserBody {
    std::string path, kernel;
    uint32_t flags = 0;
    hipModule_t mod[UB_MAX_DEVICES] = {};
    hipFunction_t fn[UB_MAX_DEVICES] = {};
};

std::mutex g_mu;
std::vector<UserBody>& bodies() {
    static std::vector<UserBody> b;
    return b;
}

// the kernel of body k on device `dev` (loaded on first use; with g_mu held)
int function_on(UserBody& b, int dev, hipFunction_t* out) {
    if (dev < 0 || dev >= UB_MAX_DEVICES) return set_error(VH_ERR_DEVICE, "device %d out of range", dev);
    if (!b.fn[dev]) {
        // a failed call leaves its error as the thread's last HIP error, which
        // a later launch check would report as its own: it is cleared here
        hipModule_t m = nullptr;
        hipError_t e = hipModuleLoad(&m, b.path.c_str());
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(VH_ERR_ARG, "can not load code object %s (%s)", b.path.c_str(), hipGetErrorString(e));
        }
        hipFunction_t f = nullptr;
        e = hipModuleGetFunction(&f, m, b.kernel.c_str());
        if (e != hipSuccess) {
            (void)hipGetLastError();
            (void)hipModuleUnload(m);
            (void)hipGetLastError();
            return set_error(VH_ERR_ARG, "no kernel %s in %s (%s)", b.kernel.c_str(), b.path.c_str(),
                             hipGetErrorString(e));
        }
        b.mod[dev] = m;
        b.fn[dev] = f;
    }
    *out = b.fn[dev];
    return VH_OK;
}

}  // namespace

extern "C" int vh_register_body(const char* code_object, const char* kernel, uint32_t flags,
                                uint32_t* transition_id) {
    if (!code_object || !kernel || !transition_id) return set_error(VH_ERR_ARG, "null argument");
    if (flags & ~(uint32_t)(VH_BODY_ADDS_EDGES | VH_BODY_ADDS_AGENTS))
        return set_error(VH_ERR_ARG, "unknown body flags");
    std::lock_guard<std::mutex> lk(g_mu);
    auto& all = bodies();
    for (size_t k = 0; k < all.size(); k++)
        if (all[k].path == code_object && all[k].kernel == kernel) {
            if (all[k].flags != flags) return set_error(VH_ERR_ARG, "body registered with other flags");
            *transition_id = VH_TR_USER + (uint32_t)k;
            return VH_OK;
        }
    if (all.size() >= VH_MAX_USER_BODIES) return set_error(VH_ERR_CAPACITY, "too many registered bodies");
    UserBody b;
    b.path = code_object;
    b.kernel = kernel;
    b.flags = flags;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipFunction_t f;
    RC(function_on(b, dev, &f));   // fail early on a missing file or kernel
    all.push_back(b);
    *transition_id = VH_TR_USER + (uint32_t)(all.size() - 1);
    return VH_OK;
}

namespace vh {

int user_body_flags(uint32_t tr, uint32_t* flags) {
    if (tr < VH_TR_USER) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    const uint32_t k = tr - VH_TR_USER;
    if (k >= bodies().size()) return -1;
    *flags = bodies()[k].flags;
    return 0;
}

int launch_user_body(uint32_t tr, const DCtx* d_ctx, int ci, int call_t, uint64_t nslots, int pass,
                     int with_edge, hipStream_t st) {
    if (!nslots) return VH_OK;
    hipFunction_t f;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        const uint32_t k = tr - VH_TR_USER;
        if (tr < VH_TR_USER || k >= bodies().size())
            return set_error(VH_ERR_UNREGISTERED, "unregistered transition %u", tr);
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        RC(function_on(bodies()[k], dev, &f));
    }
    const DCtx* c = d_ctx;
    void* args[] = {(void*)&c, (void*)&ci, (void*)&call_t, (void*)&pass, (void*)&with_edge};
    HIPCHECK(hipModuleLaunchKernel(f, grid_for(nslots, 256), 1, 1, 256, 1, 1, 0, st, args, nullptr));
    return VH_OK;
}

}  // namespace vh
