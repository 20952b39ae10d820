// rp_dyson_noalloc: the Dyson step of rp_dyson for callers that cannot
// allocate device memory (HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  The caller passes chi and eps^-1 in device memory;
// the step's scratch (status word, Drude plasma frequencies and partial sums,
// pivot indices, the Gauss-Jordan move lists) lives in one fixed file-scope
// device array, so the call makes no allocation.  It runs the same kernels as
// rp_dyson.
//
// Every GPU has its own copy of the array, at its own address: the address is
// looked up once per device (rp_init calls noalloc_init with the device
// current) and kept, and only that address is handed to kernels.  Every call
// on a device shares the array: calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "rp_internal.h"

constexpr size_t kScratchBytes = size_t(4) << 20;  // 4 MiB per device
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_mu;
std::vector<unsigned char*> g_scratch;  // device address of `scratch`, per device ordinal

unsigned char* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  return device >= 0 && device < (int)g_scratch.size() ? g_scratch[device] : nullptr;
}

inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }
}  // namespace

namespace rp {

void noalloc_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (device >= (int)g_scratch.size()) g_scratch.resize(device + 1, nullptr);
  if (g_scratch[device]) return;
  void* p = nullptr;
  RP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  RP_CHECK(p != nullptr, "rp_init: no device address for the fixed scratch");
  g_scratch[device] = (unsigned char*)p;
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

}  // namespace rp

using namespace rp;

extern "C" int rp_dyson_noalloc(rp_handle h, const double* chi_dev, double* epsinv_dev) {
  return guarded([&] {
    RP_CHECK(h && h->qready, "rp_dyson_noalloc: no q prepared");
    RP_CHECK(chi_dev && epsinv_dev, "rp_dyson_noalloc: null device pointer");
    RP_HIP(hipSetDevice(h->device));
    unsigned char* base = scratch_of(h->device);
    RP_CHECK(base, "rp_dyson_noalloc: scratch of this device not set up by rp_init");
    const int n = h->ng, nb = h->nbatch();
    const size_t nintra = h->q0 ? h->intra.size() : 0;
    // layout of the scratch: status | wd[3] | Drude partials | ipiv | move lists
    const size_t off_wd = 256;
    const size_t off_part = off_wd + align256(3 * sizeof(double));
    const size_t off_ipiv = off_part + align256(3 * nintra * sizeof(double2));
    const size_t off_work = off_ipiv + align256((size_t)nb * n * sizeof(int));
    const size_t need = off_work + align256(gauss_jordan_work_ints(nb) * sizeof(int));
    RP_CHECK(need <= kScratchBytes,
             "rp_dyson_noalloc: NG " + std::to_string(n) + " x " + std::to_string(nb) +
                 " matrices need " + std::to_string(need) + " bytes of scratch, more than its " +
                 std::to_string(kScratchBytes));
    int* status = (int*)base;
    double* wd = (double*)(base + off_wd);
    double2* part = (double2*)(base + off_part);
    int* ipiv = (int*)(base + off_ipiv);
    int* work = (int*)(base + off_work);
    hipStream_t s = h->stream;
    launch_zero(status, sizeof(int), s);  // a fill kernel: no memset node under capture
    if (nintra > 0) {
      launch_drude(h->d_coef.p, h->d_kg0.p, h->d_ng0.p, h->ntk, h->ntg, h->d_intra.p, (int)nintra,
                   h->d_sk0.p, h->d_bvec.p, h->d_eband.p, h->ef, h->delt, h->dmna, h->dmnr,
                   h->nkb[0], h->nkb[1], h->nkb[2], h->d_imt.p, h->d_index.p, h->volume, status, part,
                   wd, s);
    } else {
      launch_zero(wd, 3 * sizeof(double), s);
    }
    launch_eps_build((const double2*)chi_dev, h->d_qn.p, n, nb, h->nen, h->q0 ? 1 : 0, h->g0, wd,
                     h->d_em.p, h->delt, (double2*)epsinv_dev, s);
    launch_gauss_jordan((double2*)epsinv_dev, n, nb, ipiv, work, status, s);
    int st = 0;
    RP_HIP(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s));
    RP_HIP(hipStreamSynchronize(s));
    throw_status(st, "rp_dyson_noalloc");
  });
}
