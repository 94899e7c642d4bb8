// gamd_host.hpp -- host-side plumbing shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "gala_amd.h"
#include "gamd_device.hpp"

struct gamd_potential {
    gamd::DevPot dev;               // host copy of the table (coefficient pointers are device memory)
    const gamd::DevPot *ddev;       // the same table in device memory, read by the kernels
    std::vector<void *> buffers;    // device allocations owned by this handle
    int device;
};

namespace gamd {

#define GAMD_HIP_CHECK(expr)                                                             \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            fprintf(stderr, "gala_amd: %s failed: %s (%s:%d)\n", #expr,                  \
                    hipGetErrorString(_e), __FILE__, __LINE__);                          \
            return GAMD_ERR_HIP;                                                         \
        }                                                                                \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Launch after the caller checked shapes; surface launch errors as GAMD_ERR_HIP.
inline int launch_status() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "gala_amd: kernel launch failed: %s\n", hipGetErrorString(e));
        return GAMD_ERR_HIP;
    }
    return GAMD_OK;
}

inline unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g > 0 ? g : 1);
}

// Number of compute units of the current device (for persistent launches).
inline int cu_count() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 256;
    return cus > 0 ? cus : 256;
}

// The body potentials of an N-body system, passed to kernels by value (the
// kernel reads them from its argument segment with scalar loads).  p[j] is
// body j's device table, nullptr for a massless body.
struct BodyTable {
    const DevPot *p[GAMD_MAX_MASSIVE];
    int n;    // bodies 0..n-1 may carry a potential
    int scf;  // some body potential has an SCF component
};

inline int make_body_table(const gamd_potential *const *bodies, int nbody, BodyTable &bt) {
    if (nbody < 0 || nbody > GAMD_MAX_MASSIVE || (nbody > 0 && !bodies)) return GAMD_ERR_INVALID;
    bt = BodyTable{};
    bt.n = nbody;
    for (int j = 0; j < nbody; j++) {
        bt.p[j] = bodies[j] ? bodies[j]->ddev : nullptr;
        if (bodies[j] && bodies[j]->dev.has_scf) bt.scf = 1;
    }
    return GAMD_OK;
}

}  // namespace gamd
