// psph_noalloc.inl -- the allocation-free entry point (part of psph_device.hip's
// translation unit: it reuses that file's reduction kernel).
#include <mutex>

// largest input of the allocation-free sums: the kernels index with 32 bits
constexpr uint32_t kNoAllocMaxParticles = 0xffffff00u;

// ---------------------------------------------------------------------------
// Allocation-free conservation sums (psph_conservation_noalloc)
// ---------------------------------------------------------------------------
// The block partials live in a file-scope device array instead of a context's
// buffer: each GPU holds its own copy from module load to process end.  Host
// code never passes the array's name to a kernel (on the host it is only a
// stand-in symbol); it looks up the device address once per device and passes
// that.  Every call fully overwrites the partials it reads, so nothing needs
// clearing; calls that overlap on one GPU share the array and race.
static __device__ __align__(256) double scratch[kRedBlocks * 8];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// sums each of the 8 partial columns in block order, as psph_conservation
// does on the host, so the results are bit-identical
__global__ void k_sum_parts(const double *part, int nblocks, double *out) {
  const int k = threadIdx.x;
  if (k >= 8) return;
  double s = 0;
  for (int b = 0; b < nblocks; b++) s += part[b * 8 + k];
  out[k] = s;
}

int scratch_address(double **out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(PSPH_EINVAL, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" int psph_conservation_noalloc(const psph_soa *dev, size_t n, int32_t dim, double *out_dev,
                                         void *stream) {
  if (!dev || !out_dev) return fail(PSPH_EINVAL, "null argument");
  if (dim != 1 && dim != 3) return fail(PSPH_EINVAL, "dim must be 1 or 3");
  if (n > kNoAllocMaxParticles)
    return fail(PSPH_EINVAL, "conservation_noalloc: %zu particles exceed the maximum of %zu", n,
                (size_t)kNoAllocMaxParticles);
  const bool d3 = dim == 3;
  if (n && (!dev->bc || !dev->m || !dev->e || !dev->vx || !dev->xi || (d3 && (!dev->vy || !dev->vz || !dev->z))))
    return fail(PSPH_EINVAL, "conservation_noalloc: bc, m, e, vx, xi (and vy, vz, z in 3D) are required");
  double *part = nullptr;
  PSPH_TRY(scratch_address(&part));
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_conservation, kRedBlocks, 256, 0, s, (uint32_t)n, dev->bc, dev->m, dev->e, dev->vx,
                     d3 ? dev->vy : nullptr, d3 ? dev->vz : nullptr, dev->xi, d3 ? dev->z : nullptr, part);
  hipLaunchKernelGGL(k_sum_parts, 1, 64, 0, s, part, kRedBlocks, out_dev);
  HIP_TRY(hipGetLastError());
  return 0;
}

