// hevcr_residual_blocks_noalloc: the batched dequantisation + inverse
// transform of hevcr_residual_blocks (the reference's perform_dequant +
// perform_transform_inverse, transform.cpp:293-394, over a list of blocks)
// for callers that must not allocate device memory. Everything the kernels
// need beyond the caller's host inputs and outputs lives in one file-scope
// device array, `scratch`, of kScratchBytes per device:
//
//   coefficients in | outputs | residual plane (32 samples wide) |
//   block parameters | block records by size | row of each block |
//   counters | launch descriptor
//
// The block records and their rows are built on the device (a fill kernel
// clears the counters, no memset), the existing residual kernels run over
// them, and a gather kernel moves each block from the plane to its place in
// the outputs. Every call on a GPU uses that GPU's one array and leaves its
// contents behind; calls that overlap on one GPU (from engines on different
// host threads) race on it.
#include <cstring>
#include <mutex>

#include "runtime.hpp"

namespace hevcr {
namespace {

constexpr uint64_t kNaMaxCoeffs = 1u << 20;            // int16 coefficients per call (2 MiB)
constexpr uint32_t kNaMaxRows = 32768;                 // rows of the 32-wide residual plane (2 MiB)
constexpr uint32_t kNaMaxBlocks = kNaMaxRows / 4;      // every block takes at least 4 rows
constexpr size_t kOCoef = 0;
constexpr size_t kOOut = kOCoef + 2 * kNaMaxCoeffs;
constexpr size_t kOPlane = kOOut + 2 * kNaMaxCoeffs;
constexpr size_t kOParam = kOPlane + (size_t)kNaMaxRows * 32 * 2;  // log2, cidx, qp, flags: a byte each per block
constexpr size_t kOOff = kOParam + 4 * (size_t)kNaMaxBlocks;       // coefficient offset per block
constexpr size_t kOBlocks = kOOff + 4 * (size_t)kNaMaxBlocks;      // ResBlock records, grouped by size
constexpr size_t kORow = kOBlocks + sizeof(ResBlock) * (size_t)kNaMaxBlocks;
constexpr size_t kOCount = kORow + 4 * (size_t)kNaMaxBlocks;       // per size: records placed; [4]: rows
constexpr size_t kOMeta = (kOCount + 64 + 255) & ~(size_t)255;

struct NaMeta {
    FrameDev d;
    uint32_t pre[4][2];  // one-frame prefix arrays of the four block sizes
};
constexpr size_t kScratchBytes = kOMeta + ((sizeof(NaMeta) + 255) & ~(size_t)255);

}  // namespace

// Device memory of this entry point; host code uses the address looked up
// per device (noalloc_init), never this name.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kNaMaxDevices] = {};

__global__ void __launch_bounds__(64) k_na_clear(uint32_t* __restrict__ counters) {
    if (threadIdx.x < 8) counters[threadIdx.x] = 0;
}

// Block records: each block goes to the list of its size (slot by atomic
// count) and to the next free rows of the plane (atomic row counter).
__global__ void __launch_bounds__(256) k_na_lists(const uint8_t* __restrict__ prm, const uint32_t* __restrict__ off,
                                                  uint32_t n, uint4 base, ResBlock* __restrict__ blocks,
                                                  uint32_t* __restrict__ counters, uint32_t* __restrict__ rowmap) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t log2 = prm[i], cidx = prm[kNaMaxBlocks + i], qp = prm[2 * kNaMaxBlocks + i],
                   fl = prm[3 * kNaMaxBlocks + i];
    const uint32_t q = log2 - 2;
    const uint32_t slot = atomicAdd(&counters[q], 1u);
    const uint32_t row = atomicAdd(&counters[4], 1u << log2);
    ResBlock rb;
    rb.coeff_off = off[i];
    rb.x = 0;
    rb.y = (uint16_t)row;
    rb.log2 = (uint8_t)log2;
    rb.cidx = 0;  // every block goes to the one residual plane
    rb.qp = (uint8_t)qp;
    rb.flags = (uint8_t)(((fl & 1) ? kResIntra : 0) | ((fl & 1) && cidx == 0 && log2 == 2 ? kResDst : 0) |
                         ((fl & 2) ? kResTskip : 0));
    const uint32_t b = q == 0 ? base.x : (q == 1 ? base.y : (q == 2 ? base.z : base.w));
    blocks[b + slot] = rb;
    rowmap[i] = row;
}

// Block i's N x N residual from its plane rows to outputs + off[i], row-major.
__global__ void __launch_bounds__(256) k_na_gather(const uint8_t* __restrict__ prm, const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ rowmap, const int16_t* __restrict__ plane,
                                                   int16_t* __restrict__ out) {
    const uint32_t i = blockIdx.x;
    const int log2 = prm[i], nn = 1 << log2;
    const uint32_t row = rowmap[i], o = off[i];
    for (int k = threadIdx.x; k < nn * nn; k += 256) {
        const int y = k >> log2, x = k & (nn - 1);
        out[o + (uint32_t)k] = plane[(size_t)(row + (uint32_t)y) * 32 + (uint32_t)x];
    }
}

}  // namespace

// The device's scratch address, looked up once per device with that device
// current (hevcr_engine_create calls this).
int noalloc_init(int device, void** ptr) {
    if (device < 0 || device >= kNaMaxDevices) return fail(HEVCR_EINVAL, "device %d", device);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return fail(HEVCR_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(err));
        g_scratch[device] = p;
    }
    *ptr = g_scratch[device];
    return HEVCR_OK;
}

}  // namespace hevcr

using namespace hevcr;

extern "C" int hevcr_residual_blocks_noalloc(hevcr_engine* e, int bit_depth, uint32_t n, const uint8_t* log2,
                                             const uint8_t* cidx, const uint8_t* qp, const uint8_t* flags,
                                             const uint32_t* off, const int16_t* coeffs, uint64_t num_coeffs,
                                             int16_t* out) {
    if (!e || (n && (!log2 || !cidx || !qp || !flags || !off || !coeffs || !out)))
        return fail(HEVCR_EINVAL, "null argument");
    if (!e->scratch) return fail(HEVCR_EDEVICE, "engine without a scratch address");
    if (bit_depth < 8 || bit_depth > 10) return fail(HEVCR_EUNSUPPORTED, "bit depth %d", bit_depth);
    if (n > kNaMaxBlocks || num_coeffs > kNaMaxCoeffs)
        return fail(HEVCR_EINVAL, "%u blocks / %llu coefficients exceed the no-alloc limits (%u blocks, %llu coefficients)",
                    n, (unsigned long long)num_coeffs, kNaMaxBlocks, (unsigned long long)kNaMaxCoeffs);
    if (!n) return HEVCR_OK;
    uint32_t count[4] = {0, 0, 0, 0}, rows = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (log2[i] < 2 || log2[i] > 5 || cidx[i] > 2 || qp[i] > 51 + 6 * (bit_depth - 8) || (flags[i] & ~3u))
            return fail(HEVCR_EINVAL, "block %u parameters", i);
        if ((flags[i] & 2) && log2[i] != 2) return fail(HEVCR_EUNSUPPORTED, "block %u transform skip above 4x4", i);
        if ((uint64_t)off[i] + (1u << (2 * log2[i])) > num_coeffs) return fail(HEVCR_EINVAL, "block %u range", i);
        count[log2[i] - 2]++;
        rows += 1u << log2[i];
        if (rows > kNaMaxRows) return fail(HEVCR_EINVAL, "blocks need more than the no-alloc limit of %u rows", kNaMaxRows);
    }
    HIPCHK(hipSetDevice(e->device));
    hipStream_t s = e->stream;
    char* base = static_cast<char*>(e->scratch);
    uint8_t* prm = reinterpret_cast<uint8_t*>(base + kOParam);
    uint32_t* doff = reinterpret_cast<uint32_t*>(base + kOOff);
    ResBlock* blocks = reinterpret_cast<ResBlock*>(base + kOBlocks);
    uint32_t* rowmap = reinterpret_cast<uint32_t*>(base + kORow);
    uint32_t* counters = reinterpret_cast<uint32_t*>(base + kOCount);
    int16_t* dcoef = reinterpret_cast<int16_t*>(base + kOCoef);
    int16_t* dout = reinterpret_cast<int16_t*>(base + kOOut);
    // inputs, and the caller's outputs as they are (samples outside every block come back unchanged)
    HIPCHK(hipMemcpyAsync(dcoef, coeffs, num_coeffs * 2, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dout, out, num_coeffs * 2, hipMemcpyHostToDevice, s));
    const uint8_t* src[4] = {log2, cidx, qp, flags};
    for (int k = 0; k < 4; k++) HIPCHK(hipMemcpyAsync(prm + (size_t)k * kNaMaxBlocks, src[k], n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(doff, off, (size_t)n * 4, hipMemcpyHostToDevice, s));
    // the launch descriptor: one frame, the plane as residual target (and as the
    // prediction the kernels read under a block; no block adds to it here)
    NaMeta m;
    std::memset(&m, 0, sizeof(m));
    m.d.fc.bd = bit_depth;
    m.d.fc.maxv = (1 << bit_depth) - 1;
    m.d.fc.stride[0] = 32;
    m.d.coeffs = dcoef;
    m.d.res[0] = reinterpret_cast<int16_t*>(base + kOPlane);
    for (int c = 0; c < 3; c++) m.d.rec[c] = base + kOPlane;
    uint32_t at = 0, first[4], wgs[4];
    for (int k = 0; k < 4; k++) {
        first[k] = at;
        m.d.blocks[k] = blocks + at;
        m.d.nblocks[k] = count[k];
        at += count[k];
        wgs[k] = (count[k] + kResPerGroup[k] - 1) / (uint32_t)kResPerGroup[k];
        m.pre[k][0] = 0;
        m.pre[k][1] = wgs[k];
    }
    NaMeta* dm = reinterpret_cast<NaMeta*>(base + kOMeta);
    HIPCHK(hipMemcpyAsync(dm, &m, sizeof(m), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, s, counters);
    hipLaunchKernelGGL(k_na_lists, dim3((n + 255) / 256), dim3(256), 0, s, prm, doff, n,
                       make_uint4(first[0], first[1], first[2], first[3]), blocks, counters, rowmap);
    const uint32_t* pre[4];
    for (int k = 0; k < 4; k++) pre[k] = &dm->pre[k][0];
    launch_residual<uint16_t>(s, &dm->d, pre, 1, wgs);
    hipLaunchKernelGGL(k_na_gather, dim3(n), dim3(256), 0, s, prm, doff, rowmap,
                       reinterpret_cast<const int16_t*>(base + kOPlane), dout);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, num_coeffs * 2, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return HEVCR_OK;
}
