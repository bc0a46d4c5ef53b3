// lj_noalloc.hip -- calculateNForcesNoAlloc: the compatibility force call
// (calculateNForces) for callers that cannot allocate device memory.
//
// All of its device scratch (fixed-point positions, the force pass's slice
// partials, block rows, the RDF histogram and bin table, the pressure word)
// sits in one file-scope device array, `scratch`, of kScratchBytes; each GPU
// the library runs on holds its own copy until the process ends.  The call
// makes no hipMalloc / hipFree and creates no stream: it runs on the null
// stream and returns when the results are on the host.  Every call on a GPU
// shares that GPU's array, so calls that overlapped on one GPU would race: a
// lock serialises the calls of one process (each process has its own copy).
// The array keeps the last call's contents; the part that must start at zero
// (the RDF histogram) is cleared by a kernel on every call.
#include <mutex>
#include <vector>

#include "lj_internal.h"
#include "lj_util.h"

namespace {

using namespace lj;

// largest numBodies served
constexpr int kNoAllocMaxN = 65536;
// entries S x Npad of the sliced plan (make_force_plan(N, 0, 1)) at any
// N <= kNoAllocMaxN: at most 587520 (N = 65280: 9 slices of 65280)
constexpr size_t kPartCap = 589824;
constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t kBlkRows = (kNoAllocMaxN + kBlock - 1) / kBlock;
constexpr size_t kOffQ = 0;
constexpr size_t kOffFpart = align256(kOffQ + sizeof(uint4) * kNoAllocMaxN);
constexpr size_t kOffApart = align256(kOffFpart + sizeof(float4) * kPartCap);
constexpr size_t kOffBlk = align256(kOffApart + sizeof(float) * kPartCap);
constexpr size_t kOffRdf = align256(kOffBlk + sizeof(double) * kBlkFields * kBlkRows);
constexpr size_t kOffTab = align256(kOffRdf + sizeof(unsigned long long) * 256);
constexpr size_t kOffPressure = align256(kOffTab + sizeof(float) * 257);
constexpr size_t kScratchBytes = align256(kOffPressure + sizeof(float));  // 12.3 MiB

}  // namespace

// the call's whole device scratch (one copy per GPU)
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Per device: the address of that device's copy of `scratch` (from
// hipGetSymbolAddress with the device current; the array's name is a
// host-side stand-in and never a device address), and the RDF bin width
// whose table the copy holds.
struct NoAllocDevice {
  unsigned char* ptr = nullptr;
  float tab_dr2 = 0.f;
  float tab_last = 0.f;  // the table's last-bin threshold (host copy)
};
std::mutex g_mu;
std::vector<NoAllocDevice> g_dev;

int noalloc_device(NoAllocDevice** out)
{
  int dev = 0;
  LJ_HIP(hipGetDevice(&dev));
  if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
  NoAllocDevice& d = g_dev[dev];
  if (!d.ptr) {
    void* p = nullptr;
    LJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(LJ_ERR_DEVICE, "calculateNForcesNoAlloc: no scratch address on device %d", dev);
    d.ptr = static_cast<unsigned char*>(p);
  }
  *out = &d;
  return LJ_OK;
}

template <class T>
T* at(unsigned char* base, size_t off)
{
  return reinterpret_cast<T*>(base + off);
}

int noalloc_forces(float* Pos, float* Force, float* host_pressure, int N, float host_L,
                   int Lperiodic, int* host_RDF, float host_dr2)
{
  if (!Pos || !Force) return fail(LJ_ERR_ARG, "calculateNForcesNoAlloc: null device array");
  if (N <= 0 || N > kNoAllocMaxN)
    return fail(LJ_ERR_ARG, "calculateNForcesNoAlloc: numBodies must be in [1, %d] (got %d)",
                kNoAllocMaxN, N);
  if (Lperiodic && !(host_L > 0.f))
    return fail(LJ_ERR_ARG, "calculateNForcesNoAlloc: box side must be positive");
  if (host_RDF && !(host_dr2 > 0.f))
    return fail(LJ_ERR_ARG, "calculateNForcesNoAlloc: RDF bin width must be positive");
  LJ_TRY(check_device());
  const ForcePlan pl = make_force_plan(N, 0, 1);
  if ((size_t)pl.S * pl.Npad > kPartCap)
    return fail(LJ_ERR_ARG, "calculateNForcesNoAlloc: %d x %d partial sums exceed the scratch",
                pl.S, pl.Npad);
  const PairScalars ps = make_pair_scalars((double)host_L, Lperiodic != 0, 1., 1., true);
  NoAllocDevice* d = nullptr;
  LJ_TRY(noalloc_device(&d));
  unsigned char* base = d->ptr;
  uint4* q = at<uint4>(base, kOffQ);
  float4* fpart = at<float4>(base, kOffFpart);
  float* apart = at<float>(base, kOffApart);
  double* blk = at<double>(base, kOffBlk);
  unsigned long long* hist = at<unsigned long long>(base, kOffRdf);
  float* tab = at<float>(base, kOffTab);
  float* pressure = at<float>(base, kOffPressure);
  hipStream_t st = nullptr;  // the null stream: no stream is created
  const float4* p = reinterpret_cast<const float4*>(Pos);
  if (ps.periodic) LJ_HIP(launch_prep_q(st, p, q, N, ps.qscale));
  LJ_HIP(launch_force(st, pl, ps, q, p, fpart, apart, false, true));
  LJ_HIP(launch_reduce_forces(st, pl, ps, fpart, apart, reinterpret_cast<float4*>(Force), blk));
  LJ_HIP(launch_finalize_forces(st, blk, (N + kBlock - 1) / kBlock, 1., nullptr, pressure));
  if (host_RDF) {
    LJ_HIP(launch_fill_u64(st, hist, 256));
    if (d->tab_dr2 != host_dr2) {
      float h[257];
      rdf_bin_thresholds(LJ_RDF_REF_GPU, 256, host_dr2, 0., h);
      LJ_HIP(hipMemcpyAsync(tab, h, sizeof(h), hipMemcpyHostToDevice, st));
      LJ_HIP(hipStreamSynchronize(st));  // the host table goes out of scope
      d->tab_dr2 = host_dr2;
      d->tab_last = h[255];
    }
    RdfArgs a{};
    a.pos = p;
    a.N = N;
    a.periodic = Lperiodic != 0;
    a.L = (double)host_L;
    a.rule = LJ_RDF_REF_GPU;
    a.nbins = 256;
    a.tab = tab;
    a.est_scale = 1.f / host_dr2;
    gpu_minimage_thresholds((double)host_L, a.thr);
    a.end = d->tab_last;
    a.hist = hist;
    LJ_HIP(launch_rdf(st, a));
  }
  float pr = 0.f;
  unsigned long long h[256];
  LJ_HIP(hipMemcpyAsync(&pr, pressure, sizeof(float), hipMemcpyDeviceToHost, st));
  if (host_RDF) LJ_HIP(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, st));
  LJ_HIP(hipStreamSynchronize(st));
  if (host_pressure) *host_pressure = pr;
  if (host_RDF)
    for (int b = 0; b < 256; ++b) host_RDF[b] = (int)h[b];
  return LJ_OK;
}

}  // namespace

extern "C" {

int calculateNForcesNoAlloc(float* Pos, float* Force, float* host_pressure, int numBodies,
                            float host_L, int Lperiodic, int* host_RDF, float host_dr2,
                            int /*blockSize*/, int /*q*/)
{
  std::lock_guard<std::mutex> lk(g_mu);
  return noalloc_forces(Pos, Force, host_pressure, numBodies, host_L, Lperiodic, host_RDF,
                        host_dr2);
}

}  // extern "C"
