// ssg_noalloc.hip -- ssg_beamsearch_noalloc_dev, the form of ssg_beamsearch_dev that
// allocates nothing (no reference counterpart), for callers that cannot allocate device
// memory: code under HIP graph capture, serving loops, harnesses that forbid it.
//
// The beam kernels need scratch beyond their inputs and outputs: one visited byte per
// (wave slot, object), one epoch byte per slot and the persistent grid's work counter.
// ssg_beamsearch_dev keeps them in the handle and allocates them on first use; here they
// live in `scratch`, a file-scope device array of SCRATCH_BYTES that every GPU the library
// runs on holds until the process ends.  Its device address is resolved once per device
// (ssg_noalloc_init, called by ssg_create with the device current); host code uses only
// that address, never the array's name.
//
// Layout of the scratch: [0, 4) the work counter, [EPOCH_OFF, EPOCH_OFF + slots) the slot
// epochs, [VIS_OFF, VIS_OFF + slots * stride) the visited slots (stride = n rounded up to
// 1024).  Every call clears the part it uses with fill kernels before the search (the
// array keeps what the last call on the GPU left), so calls whose kernels can run at the
// same time on one GPU (other streams, other handles) race on it.
#include <mutex>

#include "ssg_handle.h"

namespace {

constexpr size_t SCRATCH_BYTES = size_t(256) << 20;
constexpr size_t EPOCH_OFF = 256;
constexpr size_t VIS_OFF = size_t(128) << 10;
constexpr int64_t MAX_SLOTS = (int64_t)(VIS_OFF - EPOCH_OFF);
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};

}  // namespace

int ssg_noalloc_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return ssg_fail(SSG_E_LIMIT, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return SSG_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return ssg_fail(SSG_E_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch[device] = p;
    return SSG_OK;
}

extern "C" int ssg_beamsearch_noalloc_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k,
                                          int32_t kstride, const ssg_beam_params* bp,
                                          const ssg_query_opts* qo, uint32_t* dids, float* ddists,
                                          int64_t* dcd, int64_t* dcb, void* stream)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!dQ || !dids || !ddists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (qo && qo->hints && (!qo->nhints || qo->hint_stride < 1))
        return ssg_fail(SSG_E_ARG, "per-query hints need counts and a stride");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        base = g_scratch[h->device];
    }
    if (!base) return ssg_fail(SSG_E_STATE, "no-allocation scratch not resolved for this device");
    const int64_t stride = (std::max<int64_t>(h->n, 1) + 1023) / 1024 * 1024;
    int64_t slots = (int64_t)(SCRATCH_BYTES - VIS_OFF) / stride;
    slots = std::min<int64_t>(slots, MAX_SLOTS);
    slots = std::min<int64_t>(slots, (int64_t)h->num_cu * 32);
    slots = slots / 4 * 4;
    if (slots < 4)
        return ssg_fail(SSG_E_LIMIT, "the no-allocation beam search holds databases of at most "
                                     "67076096 objects (256 MiB of scratch, 4 visited slots)");
    const BeamScratch sc{static_cast<uint8_t*>(base) + VIS_OFF, static_cast<uint8_t*>(base) + EPOCH_OFF,
                         static_cast<unsigned int*>(base), stride, (int)slots};
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    return ssg_beam_run(h, dQ, h->dim, m, k, kstride, bp, qo, dids, ddists, dcd, dcb, s, false, h->n,
                        false, &sc);
}
