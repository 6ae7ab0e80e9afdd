// hpc_corr_noalloc: the exact correlation of hpc_corr (gpuPMCC's FP64 path)
// for callers that cannot allocate device memory -- code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  The caller
// passes device observations and a device output matrix; everything else
// the kernels need (instance table, NaN flag, row map, block prefixes,
// means, standard deviations, z-scores, NaN mask) lives in one file-scope
// device array, `scratch`, of kScratchBytes.  Nothing is allocated or freed,
// no host memory is read back and every launch goes to the caller's stream.
//
// Each device holds its own copy of `scratch` for the life of the process.
// Its device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and only that address is handed to kernels: in host
// code the array's name is a host-side stand-in, not a device address.
// Every call on one device shares the array and leaves its contents behind;
// the setup kernel clears what must start at zero (NaN flag, row map) and
// the moments kernel rewrites every z-score and mask entry it later reads.
// Calls that overlap on one device (different streams, or threads without
// ordering) race on the array.
#pragma clang fp contract(off)

#include <mutex>

#include "hpc_internal.hpp"

namespace hpc {
namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr size_t kAlign = 256;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {};  // device address per device, once looked up

size_t up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

// Byte offsets of the parts of one call in the scratch array.
struct Parts {
  size_t inst, ictr, prefix, rows, mean, sd, z, nanm, end;
};

Parts parts(size_t p, size_t npad) {
  Parts q;
  q.inst = 0;
  q.ictr = q.inst + up(sizeof(Inst));
  q.prefix = q.ictr + up(sizeof(InstCounters));
  q.rows = q.prefix + up(4 * sizeof(long long));
  q.mean = q.rows + up(p * sizeof(int32_t));
  q.sd = q.mean + up(p * sizeof(double));
  q.z = q.sd + up(p * sizeof(double));
  q.nanm = q.z + up(p * npad * sizeof(double));
  q.end = q.nanm + up(p * npad);
  return q;
}

int scratch_on(int device, unsigned char **out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return hpc_fail(HPC_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                      hipGetErrorString(e));
    g_scratch[device] = static_cast<unsigned char *>(ptr);
  }
  *out = g_scratch[device];
  return HPC_OK;
}

// The one-instance table, a cleared NaN flag, the block prefixes of the
// pairwise kernels (tiled slots, then masked blocks) and the row map (every
// row belongs to instance 0).
__global__ void noalloc_setup_kernel(Inst I, Inst *inst, InstCounters *ictr, long long *prefix,
                                     int32_t *row_inst, int p, long long tiled, long long masked) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) {
    *inst = I;
    ictr->tests = ictr->tests_max = ictr->edges = ictr->removed = ictr->rec_words = 0ull;
    ictr->max_deg = 0;
    ictr->has_nan = 0;
    prefix[0] = 0;
    prefix[1] = tiled;
    prefix[2] = 0;
    prefix[3] = masked;
  }
  if (g < p) row_inst[g] = 0;
}

}  // namespace

size_t noalloc_scratch_bytes(size_t p, int n) {
  const size_t npad = ((size_t)n + 63) & ~(size_t)63;
  return parts(p, npad).end;
}

size_t noalloc_scratch_capacity() { return kScratchBytes; }

}  // namespace hpc

using namespace hpc;

extern "C" int hpc_corr_noalloc(int device, const double *x, uint64_t p, int n, double *cor,
                                void *stream) {
  if (!x || !cor) return hpc_fail(HPC_EINVAL, "observations or output is NULL");
  if (p < 2 || p > (uint64_t)kMaxP)
    return hpc_fail(HPC_EINVAL, "p = %llu outside [2, %d]", (unsigned long long)p, kMaxP);
  if (n < 1) return hpc_fail(HPC_EINVAL, "n = %d < 1", n);
  int count = 0;
  HPC_HIPCHK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count || device >= kMaxDevices)
    return hpc_fail(HPC_EINVAL, "no device %d", device);
  const size_t npad = ((size_t)n + 63) & ~(size_t)63;
  const Parts q = parts((size_t)p, npad);
  if (q.end > kScratchBytes)
    return hpc_fail(HPC_EINVAL,
                    "p = %llu, n = %d needs %zu bytes of scratch; the entry point holds %zu "
                    "(use hpc_corr)",
                    (unsigned long long)p, n, q.end, kScratchBytes);
  HPC_HIPCHK(hipSetDevice(device));
  unsigned char *base = nullptr;
  int rc = scratch_on(device, &base);
  if (rc) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);

  Inst I{};
  I.x = x;
  I.z = reinterpret_cast<double *>(base + q.z);
  I.nanm = base + q.nanm;
  I.mean = reinterpret_cast<double *>(base + q.mean);
  I.sd = reinterpret_cast<double *>(base + q.sd);
  I.cor = cor;
  I.start = 0;
  I.p = (int)p;
  I.n = n;
  I.npad = (int)npad;
  I.W = (int)((p + 63) / 64);
  I.flags = kActive;
  I.fast = 0;

  Layout L{};
  L.inst = reinterpret_cast<const Inst *>(base + q.inst);
  L.ictr = reinterpret_cast<InstCounters *>(base + q.ictr);
  L.row_inst = reinterpret_cast<const int32_t *>(base + q.rows);
  L.K = 1;
  L.p_total = (int)p;
  long long *prefix = reinterpret_cast<long long *>(base + q.prefix);
  const long long tiled = tiled_slots((int)p), masked = masked_blocks((int)p);

  hipLaunchKernelGGL(noalloc_setup_kernel, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, s, I,
                     reinterpret_cast<Inst *>(base + q.inst), L.ictr, prefix,
                     reinterpret_cast<int32_t *>(base + q.rows), (int)p, tiled, masked);
  HPC_HIPCHK(hipGetLastError());
  // means, standard deviations, z-scores and the NaN flag; then the pairwise
  // sums: the tiled kernel for NaN-free data, the masked kernel otherwise
  // (each reads the flag on the device)
  launch_moments_n(s, L, n);
  HPC_HIPCHK(hipGetLastError());
  launch_corr_sums(s, L, prefix, tiled, masked);
  HPC_HIPCHK(hipGetLastError());
  return HPC_OK;
}
