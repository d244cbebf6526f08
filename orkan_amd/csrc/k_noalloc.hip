// k_noalloc.hip -- device scratch of the allocation-free exp_diag entry point.
//
// exp_diag on a density matrix first writes the per-index phases
// (cos d_x t, sin d_x t) of all 2^n indices into scratch memory, then sweeps the
// stored entries (k_dm.hip: sincos_kernel, dm_phase_kernel).  The ordinary entry
// point takes that scratch from a library buffer that grows with hipMalloc; this
// one uses a fixed array in the library's own device image, so it allocates
// nothing: 2^ORK_NOALLOC_MAX_QUBITS complex128 (2 MiB) per device, held by every
// GPU the library runs on until the process ends.
//
// The array's name is a host-side stand-in on the host: its device address on
// the current GPU comes from hipGetSymbolAddress, taken once per device when the
// library's per-device runtime is created (ork_noalloc_setup, called by ork_dev()),
// and that address is what the kernels receive.
#include "k_common.hpp"

namespace {

constexpr int kMaxDev = 64;

// per-index phases of one exp_diag call: written in full by sincos_kernel before
// dm_phase_kernel reads them, so nothing in it has to start at zero
static __device__ __align__(256) double2 scratch[(size_t)1 << ORK_NOALLOC_MAX_QUBITS];

void *g_scratch_ptr[kMaxDev];   // device address of `scratch` on each GPU (null until set up)

}  // namespace

extern "C" int ork_noalloc_setup(void) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDev) return (int)hipErrorInvalidDevice;
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_ptr[dev] = ptr;
    return 0;
}

extern "C" int ork_exp_diag_noalloc(void *stream, int prec, int layout, int log_tile, void *data, int n,
                                    const double *diag_dev, double t) {
    if (layout == ORK_PURE) return ork_sv_exp_diag(stream, prec, data, n, diag_dev, t);
    if ((layout != ORK_PACKED && layout != ORK_TILED) || n < 1 || n > ORK_NOALLOC_MAX_QUBITS)
        return (int)hipErrorInvalidValue;
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDev || !g_scratch_ptr[dev]) return (int)hipErrorNotInitialized;
    return ork_dm_exp_diag(stream, prec, layout, log_tile, data, n, diag_dev, t, g_scratch_ptr[dev]);
}
