// rmh.cpp - host side of librmh: context, validation, plans, frame-step
// graphs and the host-memory entry points of include/rmh.h.
//
// Mirrors the reference GPU backend's contracts: typed errors and no panics on
// hostile sizes (crates/rustmpeg-vulkan/src/compute/mod.rs:13-16, 528-567),
// zero work is a no-op (compute/mod.rs:697), bounded waits
// (compute/mod.rs:416-418), results reach caller memory only after success
// (compute/deblock.rs:300-303), and a sticky failure latch
// (crates/rustmpeg-vulkan/src/accelerator.rs:38, 163-171).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cnn.h"
#include "kernels.h"
#include "lfnst_data.h"
#include "loopfilter.h"
#include "tables.h"

#define RMH_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

struct Arena {
    void* base = nullptr;
    size_t cap = 0;
    bool pinned = false;
    size_t used = 0;
    hipError_t reserve(size_t bytes) {
        used = 0;
        if (bytes <= cap) return hipSuccess;
        if (base) {
            if (pinned)
                hipHostFree(base);
            else
                hipFree(base);
            base = nullptr;
            cap = 0;
        }
        size_t want = bytes + bytes / 4 + 4096;
        hipError_t e = pinned ? hipHostMalloc(&base, want, hipHostMallocDefault) : hipMalloc(&base, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void* take(size_t bytes) {
        void* p = (char*)base + used;
        used += (bytes + 255) & ~(size_t)255;
        return p;
    }
    void release() {
        if (base) {
            if (pinned)
                hipHostFree(base);
            else
                hipFree(base);
        }
        base = nullptr;
        cap = 0;
    }
};

size_t aligned(size_t b) { return (b + 255) & ~(size_t)255; }

// A host plane into a packed device buffer (row pitch = width). Planes whose
// rows are contiguous go as one linear copy: 2-D copies from pageable host
// memory take a slow path in the runtime (about 9 ns per sample measured).
hipError_t upload_plane(void* dst, const rmh_plane& p, hipStream_t s) {
    const size_t row = (size_t)p.width * p.bytes_per_sample;
    if (p.stride == p.width) return hipMemcpyAsync(dst, p.data, row * p.height, hipMemcpyHostToDevice, s);
    return hipMemcpy2DAsync(dst, row, p.data, (size_t)p.stride * p.bytes_per_sample, row, p.height,
                            hipMemcpyHostToDevice, s);
}

}  // namespace

// ------------------------------------------------ no-allocation scratch
//
// Device memory of rmh_deblock_noalloc: the frame tables and the copies of the
// planes that the in-place deblocking filter reads. One copy per GPU, held
// until the process ends. Host code never passes `scratch` itself to a kernel
// or a copy (in host code the name is a stand-in, not a device address): it
// uses the per-device address that hipGetSymbolAddress returns.
constexpr size_t kScratchBytes = RMH_NOALLOC_SCRATCH_BYTES;
constexpr size_t kScratchTableBytes = (size_t)64 << 10;  // frame tables; the rest holds plane copies
static_assert(kScratchBytes - kScratchTableBytes == RMH_NOALLOC_PLANE_BYTES, "scratch split as rmh.h states");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on `device` (which must be current), looked up once per
// device. Returns null on failure.
void* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// Up to kWordChunk words carried in the kernel arguments: the launch copies
// them, so the host tables may go out of scope once the launch is enqueued
// (no pinned staging, no allocation).
constexpr int kWordChunk = 448;  // 1.8 KB of arguments, well inside the kernel-argument limit
struct WordChunk {
    uint32_t n;
    uint32_t w[kWordChunk];
};

__global__ __launch_bounds__(256) void write_words_kernel(uint32_t* __restrict__ dst, WordChunk c) {
    for (uint32_t i = threadIdx.x; i < c.n; i += blockDim.x) dst[i] = c.w[i];
}

// Writes `bytes` (a multiple of 4) of host data to device address dst on s.
hipError_t write_words(void* dst, const void* src, size_t bytes, hipStream_t s) {
    const uint32_t* w = (const uint32_t*)src;
    size_t n = bytes / 4;
    for (size_t at = 0; at < n; at += kWordChunk) {
        WordChunk c;
        c.n = (uint32_t)std::min<size_t>(kWordChunk, n - at);
        memcpy(c.w, w + at, c.n * 4);
        hipLaunchKernelGGL(write_words_kernel, dim3(1), dim3(256), 0, s, (uint32_t*)dst + at, c);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

struct rmh_ctx {
    int device = 0;
    uint32_t timeout_ms = 30000;
    std::atomic<bool> disabled{false};
    std::atomic<uint32_t> inject{0};
    std::mutex mu;
    std::string last_error;
    hipStream_t stream = nullptr;
    void* d_tables = nullptr;
    rmh::DeviceTables tabs{};
    Arena dev;
    Arena host;
    hipDeviceProp_t props{};
    void* scratch = nullptr;  // this device's address of `scratch`
};

enum PlanKind { kResidual = 0, kMc = 1, kDeblock = 2, kSao = 3, kEdge = 4, kLmcs = 5, kAlf = 6, kForward = 7, kGpm = 8 };

struct rmh_plan {
    rmh_ctx* ctx = nullptr;
    int kind = 0;
    // graphs that captured this plan's launches: they hold its device tables
    // and grid sizes, so the autotuner must not rewrite them meanwhile
    std::atomic<uint32_t> graph_refs{0};
    uint64_t bytes = 0;
    void* d_mem = nullptr;
    // residual
    const rmh_tu_desc* d_tus = nullptr;
    size_t n_tu = 0, n_coeff = 0;
    const uint32_t* d_order = nullptr;       // fast TUs grouped by size class, then the others
    rmh::TileItem* d_tiles = nullptr;  // tiles of the fast TUs
    size_t n_tiles = 0, n_fast = 0, tile_capacity = 0;
    uint32_t n_tiles_plain = 0;  // tiles [0, n_tiles_plain) are plain, the rest tool tiles
    uint32_t tile_lds = 0;
    // fast TUs of size class (log2 w, log2 h) are order[cls_first .. + cls_count),
    // the last cls_lfnst of them tool TUs (LFNST or dependent quantisation in
    // the kernel); tiles of the class hold residual_tile_tus >> split TUs each
    // (autotuned)
    uint32_t cls_first[7][7] = {}, cls_count[7][7] = {}, cls_lfnst[7][7] = {};
    std::vector<uint32_t> tile_breaks;  // positions in order where a transform pair starts (uniform classes)
    uint8_t split[7][7] = {};
    // recon (residual added into picture planes); null for residual output
    const rmh_tu_place* d_place = nullptr;
    const rmh_recon_frame* d_rframes = nullptr;
    const rmh::TileTu* d_tile_tus = nullptr;   // per fast TU, tile order
    const rmh::TileDst* d_tile_dst = nullptr;  // recon plans
    // prepass of the fast TUs with LFNST or dependent quantisation that the
    // tool tiles cannot apply: their indices are order[n_tu .. n_tu + n_prep), their dequantised blocks go to
    // d_prep at d_prep_off[k]
    int16_t* d_prep = nullptr;
    const uint32_t* d_prep_off = nullptr;
    size_t n_prep = 0;
    // inter
    const rmh_mc_frame* d_frames = nullptr;
    const rmh_pu_desc* d_pus = nullptr;
    const rmh::McItem* d_items = nullptr;
    const rmh_mc_bi_frame* d_bi_frames = nullptr;  // bi-prediction plans
    const rmh::McBiItem* d_bi_items = nullptr;
    size_t n_items = 0, n_items8 = 0;  // items [0, n_items8) are on 8-bit frames
    int mc_waves = 4;                   // work items per MC workgroup (rmh_mc_plan_autotune)
    // bi PUs with DMVR or BDOF: the DMVR jobs, the tiles of the refined PUs
    // and their MV pairs (4 int32 per PU, the signalled pairs until DMVR runs)
    const rmh::DmvrJob* d_dmvr = nullptr;
    size_t n_dmvr = 0;
    uint32_t dmvr_lds = 0;
    const rmh::McRefItem* d_ref_items = nullptr;
    size_t n_ref_items = 0;
    int32_t* d_mvs = nullptr;
    // GPM blend plans
    const rmh_gpm_frame* d_gpm_frames = nullptr;
    const rmh::GpmItem* d_gpm_items = nullptr;
    size_t n_gpm_items = 0;
    uint32_t gpm_scale[2] = {0, 0};
    // deblock
    const rmh_dbk_frame* d_dbk = nullptr;
    const rmh_plane* d_shadows = nullptr;  // 3 per frame: copies the tiled kernel reads
    const uint32_t* d_tile_start = nullptr;  // [n_dbk + 1] running sum of tiles, 8-bit frames first
    const uint32_t* d_frame_of = nullptr;    // [n_dbk] frame at each position of that order
    size_t n_dbk = 0;
    uint32_t total_tiles = 0, tiles8 = 0, max_rows = 0;
    bool dbk_copy = true;  // in place (shadow copies); false: d_shadows are the sources
    // edge map
    const rmh_edge_frame* d_edge_frames = nullptr;
    const uint32_t* d_cell_start = nullptr;
    size_t n_edge = 0;
    uint32_t total_cells = 0;
    // sao
    const rmh::SaoJob* d_sao_jobs = nullptr;
    const uint32_t* d_sao_block_job = nullptr;
    uint32_t n_sao_jobs = 0, sao_range[5] = {};
    // lmcs: workgroup ranges [0, 1) 8-bit, [1, 2) 16-bit planes
    const rmh::LmcsJob* d_lmcs_jobs = nullptr;
    const uint32_t* d_lmcs_block_job = nullptr;
    uint32_t lmcs_range[3] = {};
    // alf
    const rmh::AlfLumaJob* d_alf_luma = nullptr;
    const rmh::AlfChromaJob* d_alf_chroma = nullptr;
    const uint32_t* d_alf_luma_bj = nullptr;
    const uint32_t* d_alf_chroma_bj = nullptr;
    uint32_t alf_luma_range[3] = {}, alf_chroma_range[3] = {};
    // forward (encoder) plans: d_tus holds the TUs with their forward flags;
    // `inverse` is the residual plan of the same TUs that reconstructs the levels
    rmh_plan* inverse = nullptr;
};

struct rmh_graph {
    rmh_ctx* ctx = nullptr;
    std::vector<rmh_plan*> plans;  // captured plans (each holds a graph reference)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture = nullptr;
    hipStream_t side = nullptr;     // second branch of the capture (residual prepass, generic TUs)
    std::vector<hipEvent_t> events;  // fork / join events of the capture
};

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != d) hipSetDevice(d);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) hipSetDevice(prev);
    }
};

rmh_status fail(rmh_ctx* ctx, rmh_status s, const std::string& msg) {
    if (ctx) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->last_error = msg;
    }
    return s;
}

// Caller already holds ctx->mu.
rmh_status fail_locked(rmh_ctx* ctx, rmh_status s, const std::string& msg) {
    ctx->last_error = msg;
    return s;
}

rmh_status hip_fail_locked(rmh_ctx* ctx, hipError_t e, const char* what) {
    return fail_locked(ctx, RMH_EXECUTION, std::string(what) + ": " + hipGetErrorString(e));
}

// Gate shared by every device call: the failure latch and the injection hook.
rmh_status gate(rmh_ctx* ctx) {
    if (ctx->disabled.load()) return RMH_UNAVAILABLE;
    uint32_t n = ctx->inject.load();
    while (n > 0) {
        if (ctx->inject.compare_exchange_weak(n, n - 1)) return RMH_EXECUTION;
    }
    return RMH_OK;
}

// A failed or timed-out wait sets the failure latch, as Accelerator::fallback
// does on any GPU failure (crates/rustmpeg-vulkan/src/accelerator.rs:163-171):
// work still queued on the device may yet write its outputs, so the context
// takes no further calls.
rmh_status wait_stream(rmh_ctx* ctx, hipStream_t s) {
    auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(ctx->timeout_ms);
    for (;;) {
        hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) return RMH_OK;
        if (e != hipErrorNotReady) {
            ctx->disabled.store(true);
            return hip_fail_locked(ctx, e, "device execution failed");
        }
        if (std::chrono::steady_clock::now() > deadline) {
            ctx->disabled.store(true);
            return fail_locked(ctx, RMH_TIMEOUT, "device work did not finish within the timeout");
        }
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

// ---------------------------------------------------------------- tables

int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

bool matrix_ok(int type, int n) {
    bool ok;
    rmh::transform_entry(type, n, 0, 0, &ok);
    return ok;
}

void cg_log2(int lx, int ly, int* a, int* b) {  // coefficient_group_log2, vvc/transform.rs:100-178
    if (lx == 0) {
        *a = 0;
        *b = ly < 4 ? ly : 4;
    } else if (ly == 0) {
        *a = lx < 4 ? lx : 4;
        *b = 0;
    } else if (lx == 1) {
        *a = 1;
        *b = ly <= 2 ? 1 : 3;
    } else if (ly == 1) {
        *a = lx == 2 ? 1 : 3;
        *b = 1;
    } else {
        *a = 2;
        *b = 2;
    }
}

void diag(int w, int h, std::vector<int>& xs, std::vector<int>& ys) {  // vvc/transform.rs:189-212
    int line = 0, col = 0;
    for (int i = 0; i < w * h; i++) {
        xs.push_back(col);
        ys.push_back(line);
        if (col == w - 1 || line == 0) {
            line += col + 1;
            col = 0;
            if (line >= h) {
                col += line - (h - 1);
                line = h - 1;
            }
        } else {
            col++;
            line--;
        }
    }
}

std::vector<uint16_t> coeff_scan(int lw, int lh) {  // CoeffScan::new, vvc/transform.rs:37-79
    int a, b;
    cg_log2(lw, lh, &a, &b);
    int w = 1 << lw, h = 1 << lh;
    int gw = (w < 32 ? w : 32) >> a, gh = (h < 32 ? h : 32) >> b;
    std::vector<int> gx, gy, ix, iy;
    diag(gw, gh, gx, gy);
    diag(1 << a, 1 << b, ix, iy);
    std::vector<uint16_t> r;
    for (size_t g = 0; g < gx.size(); g++)
        for (size_t i = 0; i < ix.size(); i++) r.push_back((uint16_t)((gy[g] * (1 << b) + iy[i]) * w + gx[g] * (1 << a) + ix[i]));
    return r;
}

hipError_t build_tables(rmh_ctx* ctx) {
    std::vector<uint32_t> mat;
    for (int t = 0; t < 3; t++)
        for (int l = 0; l < 7; l++) {
            int n = 1 << l;
            ctx->tabs.mat_off[t][l] = -1;
            if (n < 2 || !matrix_ok(t, n)) continue;
            ctx->tabs.mat_off[t][l] = (int32_t)mat.size();
            bool ok;
            for (int kk = 0; kk < n / 2; kk++)
                for (int j = 0; j < n; j++) {
                    uint32_t lo = (uint16_t)(int16_t)rmh::transform_entry(t, n, 2 * kk, j, &ok);
                    uint32_t hi = (uint16_t)(int16_t)rmh::transform_entry(t, n, 2 * kk + 1, j, &ok);
                    mat.push_back(lo | (hi << 16));
                }
        }
    std::vector<uint32_t> col16;
    for (int t = 0; t < 3; t++)
        for (int l = 0; l < 7; l++) {
            int n = 1 << l;
            ctx->tabs.col16_off[t][l] = -1;
            if (n < 2 || !matrix_ok(t, n)) continue;
            ctx->tabs.col16_off[t][l] = (int32_t)col16.size();
            bool ok;
            for (int j = 0; j < n; j++)
                for (int kk = 0; kk < 16; kk++) {
                    uint32_t lo = 0, hi = 0;
                    if (2 * kk < n) lo = (uint16_t)(int16_t)rmh::transform_entry(t, n, 2 * kk, j, &ok);
                    if (2 * kk + 1 < n) hi = (uint16_t)(int16_t)rmh::transform_entry(t, n, 2 * kk + 1, j, &ok);
                    col16.push_back(lo | (hi << 16));
                }
        }
    // the even/odd split of DCT-II, for the butterflies of the uniform
    // residual passes (their output pairs j, n - 1 - j use the symmetry
    // M[k][n - 1 - j] = (-1)^k M[k][j] of these matrices)
    for (int l = 0; l < 7; l++) {
        const int n = 1 << l;
        ctx->tabs.eo_off[l] = -1;
        if (l < 3 || l > 5) continue;
        ctx->tabs.eo_off[l] = (int32_t)col16.size();
        bool ok;
        for (int j = 0; j < n; j++)
            for (int odd = 0; odd < 2; odd++)
                for (int q = 0; q < 8; q++) {
                    uint32_t lo = 0, hi = 0;
                    if (4 * q + 3 < n) {
                        lo = (uint16_t)(int16_t)rmh::transform_entry(RMH_TR_DCT2, n, 4 * q + odd, j, &ok);
                        hi = (uint16_t)(int16_t)rmh::transform_entry(RMH_TR_DCT2, n, 4 * q + 2 + odd, j, &ok);
                    }
                    col16.push_back(lo | (hi << 16));
                }
    }
    std::vector<uint16_t> scans;
    for (int a = 0; a < 7; a++)
        for (int b = 0; b < 7; b++) {
            ctx->tabs.scan_off[a][b] = (int32_t)scans.size();
            ctx->tabs.scan_len[a][b] = 0;
            if (a + b == 0) continue;
            std::vector<uint16_t> s = coeff_scan(a, b);
            ctx->tabs.scan_len[a][b] = (int32_t)s.size();
            scans.insert(scans.end(), s.begin(), s.end());
        }
    size_t o_mat = 0, o_l8 = aligned(mat.size() * 4), o_l4 = o_l8 + aligned(sizeof(RMH_LFNST8));
    size_t o_sc = o_l4 + aligned(sizeof(RMH_LFNST4)), o_c16 = o_sc + aligned(scans.size() * 2);
    size_t total = o_c16 + aligned(col16.size() * 4);
    std::vector<uint8_t> blob(total, 0);
    memcpy(blob.data() + o_mat, mat.data(), mat.size() * 4);
    memcpy(blob.data() + o_l8, RMH_LFNST8, sizeof(RMH_LFNST8));
    memcpy(blob.data() + o_l4, RMH_LFNST4, sizeof(RMH_LFNST4));
    memcpy(blob.data() + o_sc, scans.data(), scans.size() * 2);
    memcpy(blob.data() + o_c16, col16.data(), col16.size() * 4);
    hipError_t e = hipMalloc(&ctx->d_tables, total);
    if (e != hipSuccess) return e;
    e = hipMemcpy(ctx->d_tables, blob.data(), total, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    char* base = (char*)ctx->d_tables;
    ctx->tabs.mat = (const uint32_t*)(base + o_mat);
    ctx->tabs.lfnst8 = (const int8_t*)(base + o_l8);
    ctx->tabs.lfnst4 = (const int8_t*)(base + o_l4);
    ctx->tabs.scans = (const uint16_t*)(base + o_sc);
    ctx->tabs.col16 = (const uint32_t*)(base + o_c16);
    return hipSuccess;
}

// ------------------------------------------------------------ validation

// Checks one TU against the contract of inverse_residual / residual_of.
const char* check_tu(const rmh_tu_desc& t, size_t n_coeff) {
    if (t.log2_w > 6 || t.log2_h > 6) return "transform size above 64";
    const int w = 1 << t.log2_w, h = 1 << t.log2_h;
    if (w * h < 2) return "transform block smaller than 2 samples";
    if (t.bit_depth < 8 || t.bit_depth > 12) return "bit depth outside 8..12";
    if (t.qp > 63 + 6 * 4) return "QP' above 87";
    if (t.flags & ~(RMH_TU_TRANSFORM_SKIP | RMH_TU_DEP_QUANT)) return "unknown TU flags";
    if ((t.joint & ~7) || (t.joint && (t.joint & 3) == 0)) return "joint Cb-Cr code outside modes 1..3 and the sign bit";
    if ((uint64_t)t.offset + (uint64_t)(w * h) > n_coeff) return "TU block outside the coefficient buffer";
    if (t.flags & RMH_TU_TRANSFORM_SKIP) return nullptr;
    if (t.tr_hor > 2 || t.tr_ver > 2) return "unknown transform type";
    if (w == 1) {
        if (!matrix_ok(t.tr_ver, h)) return "unsupported vertical transform";
    } else if (h == 1) {
        if (!matrix_ok(t.tr_hor, w)) return "unsupported horizontal transform";
    } else if (!matrix_ok(t.tr_hor, w) || !matrix_ok(t.tr_ver, h)) {
        return "unsupported transform type for the block size";
    }
    if (t.lfnst_idx > 2) return "LFNST index outside 0..2";
    if (t.lfnst_idx) {
        if (w < 4 || h < 4) return "LFNST requires width and height >= 4";
        if (t.lfnst_mode > 94) return "LFNST intra mode outside 0..94";
    }
    int sw = (t.tr_hor != RMH_TR_DCT2 && w == 32) ? 16 : (w > 32 ? w - 32 : 0);
    int sh = (t.tr_ver != RMH_TR_DCT2 && h == 32) ? 16 : (h > 32 ? h - 32 : 0);
    sw = t.skip_w > sw ? t.skip_w : sw;
    sh = t.skip_h > sh ? t.skip_h : sh;
    if (t.lfnst_idx) {
        int nz = (w == 4 || h == 4) ? 4 : 8;
        sw = w - nz > sw ? w - nz : sw;
        sh = h - nz > sh ? h - nz : sh;
    }
    if ((w > 1 && sw >= w) || (h > 1 && sh >= h)) return "transform zero-out leaves no coefficients";
    return nullptr;
}

rmh_status validate_tus(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu, size_t n_coeff, bool locked) {
    if (n_tu > 0 && !tus) return (locked ? fail_locked : fail)(ctx, RMH_INVALID_INPUT, "null TU array");
    if (n_tu > 0x7fffffffu) return (locked ? fail_locked : fail)(ctx, RMH_LIMIT_EXCEEDED, "more than 2^31-1 TUs");
    if (n_coeff > 0xffffffffull) return (locked ? fail_locked : fail)(ctx, RMH_LIMIT_EXCEEDED, "more than 2^32 coefficients");
    for (size_t i = 0; i < n_tu; i++) {
        const char* why = check_tu(tus[i], n_coeff);
        if (why) {
            char buf[160];
            snprintf(buf, sizeof buf, "TU %zu: %s", i, why);
            return (locked ? fail_locked : fail)(ctx, RMH_INVALID_INPUT, buf);
        }
    }
    return RMH_OK;
}

bool plane_ok(const rmh_plane& p);

// Recon placement: every TU block lies inside its plane, has the plane's bit
// depth, and no two blocks of one plane overlap (the in-place add would race).
rmh_status validate_place(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu, const rmh_tu_place* place,
                          const rmh_recon_frame* frames, size_t n_frames) {
    if (n_tu && !place) return fail(ctx, RMH_INVALID_INPUT, "null placement array");
    if (n_frames && !frames) return fail(ctx, RMH_INVALID_INPUT, "null frame array");
    if (n_frames > 65535) return fail(ctx, RMH_LIMIT_EXCEEDED, "more than 65535 frames");
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++)
            if (!plane_ok(frames[f].planes[c])) return fail(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": invalid plane");
    std::vector<std::vector<uint32_t>> by_plane(n_frames * 3);
    for (size_t i = 0; i < n_tu; i++) {
        const rmh_tu_place& pl = place[i];
        if (pl.frame >= n_frames || pl.plane > 2 || pl.reserved)
            return fail(ctx, RMH_INVALID_INPUT, "TU " + std::to_string(i) + ": invalid placement");
        const rmh_plane& p = frames[pl.frame].planes[pl.plane];
        const uint32_t w = 1u << tus[i].log2_w, h = 1u << tus[i].log2_h;
        if ((uint32_t)pl.x + w > p.width || (uint32_t)pl.y + h > p.height)
            return fail(ctx, RMH_INVALID_INPUT, "TU " + std::to_string(i) + ": block outside its plane");
        if (tus[i].bit_depth != p.bit_depth)
            return fail(ctx, RMH_INVALID_INPUT, "TU " + std::to_string(i) + ": bit depth differs from its plane");
        by_plane[pl.frame * 3 + pl.plane].push_back((uint32_t)i);
        if (tus[i].joint) {
            // the coded component is Cr for mode 1 and Cb for modes 2 and 3; the
            // derived residual lands in the same block of the other chroma plane
            const int coded = (tus[i].joint & 3) == 1 ? 2 : 1;
            const rmh_plane& q = frames[pl.frame].planes[3 - coded];
            if (pl.plane != coded || q.width != p.width || q.height != p.height || q.bit_depth != p.bit_depth)
                return fail(ctx, RMH_INVALID_INPUT, "TU " + std::to_string(i) + ": joint Cb-Cr placement");
            by_plane[pl.frame * 3 + 3 - coded].push_back((uint32_t)i);
        }
    }
    std::vector<uint8_t> cover;
    for (size_t k = 0; k < by_plane.size(); k++) {
        if (by_plane[k].empty()) continue;
        const rmh_plane& p = frames[k / 3].planes[k % 3];
        cover.assign((size_t)p.width * p.height, 0);
        for (uint32_t i : by_plane[k]) {
            const uint32_t w = 1u << tus[i].log2_w, h = 1u << tus[i].log2_h;
            for (uint32_t y = 0; y < h; y++) {
                uint8_t* row = cover.data() + (size_t)(place[i].y + y) * p.width + place[i].x;
                for (uint32_t x = 0; x < w; x++) {
                    if (row[x]) return fail(ctx, RMH_INVALID_INPUT, "TU " + std::to_string(i) + ": block overlaps another TU");
                    row[x] = 1;
                }
            }
        }
    }
    return RMH_OK;
}

uint64_t residual_bytes(const rmh_tu_desc* tus, size_t n_tu) {
    uint64_t s = 0;
    for (size_t i = 0; i < n_tu; i++) s += (uint64_t)1 << (tus[i].log2_w + tus[i].log2_h);
    return 4 * s + 16 * (uint64_t)n_tu;
}

// Planes up to 65535 x 65535 samples whose rows span less than 4 GiB
// (kernels address samples with 32-bit offsets).
bool plane_ok(const rmh_plane& p) {
    return p.data && p.width > 0 && p.height > 0 && p.stride >= p.width && (p.bytes_per_sample == 1 || p.bytes_per_sample == 2) &&
           p.bit_depth >= 8 && p.bit_depth <= 12 && (p.bytes_per_sample == 2 || p.bit_depth == 8) && p.width <= 65535 &&
           p.height <= 65535 && (uint64_t)p.stride * p.height * p.bytes_per_sample <= 0xffffffffull;
}

// Address range a plane's samples occupy, and whether the batch writes it.
struct PlaneSpan {
    uintptr_t lo, hi;
    bool written;
};

void add_span(std::vector<PlaneSpan>& v, const rmh_plane& p, bool written) {
    const uintptr_t lo = (uintptr_t)p.data;
    const uint64_t n = ((uint64_t)(p.height - 1) * p.stride + p.width) * p.bytes_per_sample;
    v.push_back({lo, lo + (uintptr_t)n, written});
}

// True when a written plane overlaps any other plane of the batch (read-only
// planes may share memory). One sort and a sweep: if a written span overlaps
// an earlier one, the earlier span reaching furthest overlaps it too, or
// overlapped it when it arrived.
bool written_overlap(std::vector<PlaneSpan>& v) {
    std::sort(v.begin(), v.end(), [](const PlaneSpan& a, const PlaneSpan& b) { return a.lo < b.lo; });
    bool have = false;
    PlaneSpan far{};
    for (const PlaneSpan& s : v) {
        if (have && s.lo < far.hi && (s.written || far.written)) return true;
        if (!have || s.hi > far.hi) far = s;
        have = true;
    }
    return false;
}

rmh_status validate_mc(rmh_ctx* ctx, const rmh_mc_frame* frames, size_t n_frames, const rmh_pu_desc* pus, size_t n_pu) {
    if ((n_frames && !frames) || (n_pu && !pus)) return fail_locked(ctx, RMH_INVALID_INPUT, "null frame or PU array");
    if (n_frames > 65536) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, "more than 65536 frames");
    for (size_t f = 0; f < n_frames; f++) {
        const rmh_mc_frame& fr = frames[f];
        for (int c = 0; c < 3; c++)
            if (!plane_ok(fr.ref[c]) || !plane_ok(fr.dst[c]))
                return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": invalid plane");
        for (int c = 0; c < 3; c++) {
            if (fr.ref[c].bit_depth != fr.ref[0].bit_depth || fr.dst[c].bit_depth != fr.ref[0].bit_depth ||
                fr.dst[c].bytes_per_sample != fr.ref[c].bytes_per_sample)
                return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": mixed sample formats");
            if (fr.dst[c].width != fr.ref[c].width || fr.dst[c].height != fr.ref[c].height)
                return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": ref/dst size mismatch");
        }
        if (fr.ref[1].width != (fr.ref[0].width + 1) / 2 || fr.ref[1].height != (fr.ref[0].height + 1) / 2 ||
            fr.ref[2].width != fr.ref[1].width || fr.ref[2].height != fr.ref[1].height)
            return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": chroma planes are not 4:2:0");
    }
    std::vector<PlaneSpan> spans;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            add_span(spans, frames[f].ref[c], false);
            add_span(spans, frames[f].dst[c], true);
        }
    if (written_overlap(spans)) return fail_locked(ctx, RMH_INVALID_INPUT, "an output plane overlaps another plane of the batch");
    for (size_t i = 0; i < n_pu; i++) {
        const rmh_pu_desc& p = pus[i];
        int lw = p.log2_w & 15, lh = p.log2_h & 15;
        if ((p.log2_w & ~(15 | RMH_PU_ALT_HPEL)) || lw < 3 || lw > 7 || lh < 3 || lh > 7)
            return fail_locked(ctx, RMH_INVALID_INPUT, "PU " + std::to_string(i) + ": size outside 8..128");
        if (p.frame >= n_frames) return fail_locked(ctx, RMH_INVALID_INPUT, "PU " + std::to_string(i) + ": frame index");
        const rmh_plane& y = frames[p.frame].ref[0];
        if (p.x >= y.width || p.y >= y.height)
            return fail_locked(ctx, RMH_INVALID_INPUT, "PU " + std::to_string(i) + ": outside the picture");
        if (p.mv_x < -(1 << 17) || p.mv_x >= (1 << 17) || p.mv_y < -(1 << 17) || p.mv_y >= (1 << 17))
            return fail_locked(ctx, RMH_INVALID_INPUT, "PU " + std::to_string(i) + ": MV outside 18-bit storage");
    }
    return RMH_OK;
}

rmh_status validate_mc_bi(rmh_ctx* ctx, const rmh_mc_bi_frame* frames, size_t n_frames, const rmh_bi_pu_desc* pus,
                          size_t n_pu) {
    if ((n_frames && !frames) || (n_pu && !pus)) return fail_locked(ctx, RMH_INVALID_INPUT, "null frame or PU array");
    if (n_frames > 65536) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, "more than 65536 frames");
    for (size_t f = 0; f < n_frames; f++) {
        const rmh_mc_bi_frame& fr = frames[f];
        const std::string tag = "frame " + std::to_string(f) + ": ";
        for (int c = 0; c < 3; c++) {
            if (!plane_ok(fr.ref0[c]) || !plane_ok(fr.ref1[c]) || !plane_ok(fr.dst[c]))
                return fail_locked(ctx, RMH_INVALID_INPUT, tag + "invalid plane");
            for (const rmh_plane* q : {&fr.ref1[c], &fr.dst[c]}) {
                if (q->bit_depth != fr.ref0[0].bit_depth || q->bytes_per_sample != fr.ref0[c].bytes_per_sample ||
                    fr.ref0[c].bit_depth != fr.ref0[0].bit_depth)
                    return fail_locked(ctx, RMH_INVALID_INPUT, tag + "mixed sample formats");
                if (q->width != fr.ref0[c].width || q->height != fr.ref0[c].height)
                    return fail_locked(ctx, RMH_INVALID_INPUT, tag + "reference/output size mismatch");
            }
        }
        if (fr.ref0[1].width != (fr.ref0[0].width + 1) / 2 || fr.ref0[1].height != (fr.ref0[0].height + 1) / 2 ||
            fr.ref0[2].width != fr.ref0[1].width || fr.ref0[2].height != fr.ref0[1].height)
            return fail_locked(ctx, RMH_INVALID_INPUT, tag + "chroma planes are not 4:2:0");
    }
    std::vector<PlaneSpan> spans;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            add_span(spans, frames[f].ref0[c], false);
            add_span(spans, frames[f].ref1[c], false);
            add_span(spans, frames[f].dst[c], true);
        }
    if (written_overlap(spans)) return fail_locked(ctx, RMH_INVALID_INPUT, "an output plane overlaps another plane of the batch");
    for (size_t i = 0; i < n_pu; i++) {
        const rmh_bi_pu_desc& p = pus[i];
        const std::string tag = "PU " + std::to_string(i) + ": ";
        int lw = p.log2_w & 15, lh = p.log2_h & 15;
        if ((p.log2_w & ~(15 | RMH_PU_ALT_HPEL)) || lw < 3 || lw > 7 || lh < 3 || lh > 7)
            return fail_locked(ctx, RMH_INVALID_INPUT, tag + "size outside 8..128");
        if (p.frame >= n_frames) return fail_locked(ctx, RMH_INVALID_INPUT, tag + "frame index");
        const rmh_plane& y = frames[p.frame].ref0[0];
        if (p.x >= y.width || p.y >= y.height) return fail_locked(ctx, RMH_INVALID_INPUT, tag + "outside the picture");
        for (int32_t m : {p.mv0_x, p.mv0_y, p.mv1_x, p.mv1_y})
            if (m < -(1 << 17) || m >= (1 << 17)) return fail_locked(ctx, RMH_INVALID_INPUT, tag + "MV outside 18-bit storage");
        if (p.reserved[0] | p.reserved[1])
            return fail_locked(ctx, RMH_INVALID_INPUT, tag + "reserved bytes must be zero");
        if (p.refine & ~(RMH_BI_DMVR | RMH_BI_BDOF))
            return fail_locked(ctx, RMH_INVALID_INPUT, tag + "unknown refinement flags");
        if (p.refine && frames[p.frame].ref0[0].bit_depth != 8)
            return fail_locked(ctx, RMH_INVALID_INPUT, tag + "DMVR / BDOF are defined for 8-bit frames only");
    }
    return RMH_OK;
}

rmh_status validate_dbk(rmh_ctx* ctx, const rmh_dbk_frame* frames, size_t n_frames, bool in_place = true) {
    if (n_frames && !frames) return fail_locked(ctx, RMH_INVALID_INPUT, "null frame array");
    if (n_frames > 65535) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, "more than 65535 frames");
    for (size_t f = 0; f < n_frames; f++) {
        const rmh_dbk_frame& d = frames[f];
        for (int c = 0; c < 3; c++)
            if (!plane_ok(d.planes[c]) || d.planes[c].bit_depth != d.planes[0].bit_depth)
                return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": invalid plane");
        if (!d.edges) return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": null edge map");
        if (d.grid_w != (d.planes[0].width + 3) / 4 || d.grid_h != (d.planes[0].height + 3) / 4)
            return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": edge grid size");
        if (d.planes[1].width != (d.planes[0].width + 1) / 2 || d.planes[1].height != (d.planes[0].height + 1) / 2 ||
            d.planes[2].width != d.planes[1].width || d.planes[2].height != d.planes[1].height)
            return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": chroma planes are not 4:2:0");
    }
    if (!in_place) return RMH_OK;
    // in place: every plane is written
    std::vector<PlaneSpan> spans;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) add_span(spans, frames[f].planes[c], true);
    if (written_overlap(spans)) return fail_locked(ctx, RMH_INVALID_INPUT, "planes of the batch overlap");
    return RMH_OK;
}

// The SAO kernel reads a parameter record as 5 dwords.
static_assert(sizeof(rmh_sao_param) == 20 && offsetof(rmh_sao_param, offsets) == 4, "rmh_sao_param layout");
static_assert(sizeof(rmh_lmcs_aps) == 36 && sizeof(rmh_lmcs_frame) == 88, "rmh_lmcs_* layout");
static_assert(sizeof(rmh_alf_ctb) == 8 && sizeof(rmh_alf_aps) == 1186 && offsetof(rmh_alf_aps, chroma_coeff) == 930 &&
                  sizeof(rmh_alf_frame) == 208,
              "rmh_alf_* layout");

// SAO frames: matching src/dst planes of one bit depth, 4:2:0 chroma, a CTB
// grid that covers the visible region, src and dst not overlapping. With
// host_params, the parameters themselves are checked too (plans hold device
// addresses, where an unknown type acts as RMH_SAO_OFF).
rmh_status validate_sao(rmh_ctx* ctx, const rmh_sao_frame* frames, size_t n_frames, bool host_params) {
    if (n_frames && !frames) return fail_locked(ctx, RMH_INVALID_INPUT, "null frame array");
    if (n_frames > 65535) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, "more than 65535 frames");
    for (size_t f = 0; f < n_frames; f++) {
        const rmh_sao_frame& d = frames[f];
        const std::string tag = "frame " + std::to_string(f) + ": ";
        for (int c = 0; c < 3; c++) {
            const rmh_plane &a = d.src[c], &b = d.dst[c];
            if (!plane_ok(a) || !plane_ok(b) || a.width != b.width || a.height != b.height ||
                a.bytes_per_sample != b.bytes_per_sample || a.stride != b.stride || a.bit_depth != d.src[0].bit_depth ||
                b.bit_depth != a.bit_depth)
                return fail_locked(ctx, RMH_INVALID_INPUT, tag + "invalid or mismatched src/dst plane");
            const uint64_t n = (uint64_t)a.stride * a.height * a.bytes_per_sample;
            const uintptr_t sa = (uintptr_t)a.data, da = (uintptr_t)b.data;
            if (sa < da + n && da < sa + n) return fail_locked(ctx, RMH_INVALID_INPUT, tag + "dst overlaps src");
        }
        if (d.src[1].width != (d.src[0].width + 1) / 2 || d.src[1].height != (d.src[0].height + 1) / 2 ||
            d.src[2].width != d.src[1].width || d.src[2].height != d.src[1].height)
            return fail_locked(ctx, RMH_INVALID_INPUT, tag + "chroma planes are not 4:2:0");
        if (d.ctb_size != 16 && d.ctb_size != 32 && d.ctb_size != 64 && d.ctb_size != 128)
            return fail_locked(ctx, RMH_INVALID_INPUT, tag + "CTB size");
        if (d.vis_w == 0 || d.vis_h == 0 || d.vis_w > d.src[0].width || d.vis_h > d.src[0].height)
            return fail_locked(ctx, RMH_INVALID_INPUT, tag + "visible size");
        if (!d.params || d.ctb_cols < (d.vis_w + d.ctb_size - 1) / d.ctb_size ||
            d.ctb_rows < (d.vis_h + d.ctb_size - 1) / d.ctb_size)
            return fail_locked(ctx, RMH_INVALID_INPUT, tag + "CTB parameter grid");
        if (host_params) {
            const size_t np = (size_t)d.ctb_cols * d.ctb_rows * 3;
            for (size_t i = 0; i < np; i++) {
                const rmh_sao_param& q = d.params[i];
                if (q.type > RMH_SAO_EDGE) return fail_locked(ctx, RMH_INVALID_INPUT, tag + "SAO type");
                for (int k = 0; k < 4; k++)
                    if (q.offsets[k] < -65535 || q.offsets[k] > 65535)
                        return fail_locked(ctx, RMH_INVALID_INPUT, tag + "SAO offset out of range");
            }
        }
    }
    std::vector<PlaneSpan> spans;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            add_span(spans, frames[f].src[c], false);
            add_span(spans, frames[f].dst[c], true);
        }
    if (written_overlap(spans)) return fail_locked(ctx, RMH_INVALID_INPUT, "a dst plane overlaps another plane of the batch");
    return RMH_OK;
}

// One job per plane, planes of 8-bit storage first; params[f] is frame f's
// parameter address as the kernel will read it.
struct SaoJobs {
    std::vector<rmh::SaoJob> jobs;
    std::vector<uint32_t> block_job;  // job of each workgroup
    uint32_t range[5] = {};           // workgroup ranges of the four job groups (launch_sao)
};

SaoJobs sao_jobs(const rmh_sao_frame* frames, size_t n_frames, const std::vector<const rmh_sao_param*>& params,
                 const std::vector<std::array<rmh_plane, 6>>& planes) {
    SaoJobs out;
    // groups: row form 8-bit, CTB form 8-bit, row form 16-bit, CTB form 16-bit
    for (int group = 0; group < 4; group++) {
        for (size_t f = 0; f < n_frames; f++)
            for (int c = 0; c < 3; c++) {
                const rmh_plane& src = planes[f][c];
                const rmh_sao_frame& fr = frames[f];
                const uint32_t cs = c ? fr.ctb_size / 2 : fr.ctb_size;
                const bool ctb_form = cs >= 16;
                if ((src.bytes_per_sample == 1) != (group < 2) || ctb_form != (group % 2 == 1)) continue;
                rmh::SaoJob j{};
                j.src = src.data;
                j.dst = planes[f][3 + c].data;
                j.params = params[f];
                j.width = src.width;
                j.height = src.height;
                j.stride = src.stride;
                j.vis_w = c ? (fr.vis_w + 1) / 2 : fr.vis_w;
                j.vis_h = c ? (fr.vis_h + 1) / 2 : fr.vis_h;
                j.ctb = cs;
                j.ctb_log2 = (uint32_t)__builtin_ctz(cs);
                j.ctb_cols = fr.ctb_cols;
                j.comp = (uint32_t)c;
                j.bd = src.bit_depth;
                j.first_block = (uint32_t)out.block_job.size();
                j.vec = ((uintptr_t)src.data % 8 == 0 && (uintptr_t)j.dst % 8 == 0 &&
                         ((uint64_t)src.stride * src.bytes_per_sample) % 8 == 0);
                uint32_t nb;
                if (ctb_form) {
                    // a wave's block: la lanes of 4 samples across, 64 / la rows
                    const uint32_t la = std::min<uint32_t>(16, cs / 4), tw = 4 * la, th = 64 / la;
                    j.lanes_across = la;
                    j.tiles_x = (src.width + tw - 1) / tw;
                    j.tiles = j.tiles_x * ((src.height + th - 1) / th);
                    j.inv_tiles_x = 1.0 / (double)j.tiles_x;
                    nb = (j.tiles + 4 * rmh::kSaoBlocksPerWave - 1) / (4 * rmh::kSaoBlocksPerWave);
                } else {
                    const uint32_t k = 8 / src.bytes_per_sample;  // samples per lane
                    j.blocks_per_row = ((src.width + k - 1) / k + 255) / 256;
                    nb = j.blocks_per_row * ((src.height + rmh::kSaoRows - 1) / rmh::kSaoRows);
                }
                out.block_job.insert(out.block_job.end(), nb, (uint32_t)out.jobs.size());
                out.jobs.push_back(j);
            }
        out.range[group + 1] = (uint32_t)out.block_job.size();
    }
    return out;
}

// 16x16 luma tiles of every PU (PU order, raster inside a PU), clipped to the
// picture; tiles wholly outside it are dropped.
// The work items of one PU: its 16x16 luma tiles clipped to the picture
// (dst: the frame's three output planes), each with its chroma tiles;
// emit(x0, y0, tw, th, cw, ch) per kept tile.
template <typename F>
void pu_tiles(const rmh_plane* dst, int px, int py, int w, int h, F&& emit) {
    for (int ty = 0; ty < (h + 15) / 16; ty++)
        for (int tx = 0; tx < (w + 15) / 16; tx++) {
            const int x0 = px + 16 * tx, y0 = py + 16 * ty;
            const int tw = std::max(0, std::min({16, w - 16 * tx, (int)dst[0].width - x0}));
            const int th = std::max(0, std::min({16, h - 16 * ty, (int)dst[0].height - y0}));
            const int cw = std::max(0, std::min(std::min(16, w - 16 * tx) / 2, (int)dst[1].width - x0 / 2));
            const int ch = std::max(0, std::min(std::min(16, h - 16 * ty) / 2, (int)dst[1].height - y0 / 2));
            if ((tw == 0 || th == 0) && (cw == 0 || ch == 0)) continue;
            emit(x0, y0, tw, th, cw, ch);  // a kept tile has x0 <= width and y0 <= height, so they fit 16 bits
        }
}

std::vector<rmh::McItem> mc_items(const rmh_mc_frame* frames, const rmh_pu_desc* pus, size_t n_pu) {
    std::vector<rmh::McItem> items;
    for (size_t i = 0; i < n_pu; i++) {
        const rmh_pu_desc& pu = pus[i];
        pu_tiles(frames[pu.frame].dst, pu.x, pu.y, 1 << (pu.log2_w & 15), 1 << (pu.log2_h & 15),
                 [&](int x0, int y0, int tw, int th, int cw, int ch) {
                     rmh::McItem it{};
                     it.x0 = (uint16_t)x0;
                     it.y0 = (uint16_t)y0;
                     it.tw = (uint8_t)tw;
                     it.th = (uint8_t)th;
                     it.cw = (uint8_t)cw;
                     it.ch = (uint8_t)ch;
                     it.frame = pu.frame;
                     it.flags = (pu.log2_w & RMH_PU_ALT_HPEL) ? 1 : 0;
                     it.mv_x = pu.mv_x;
                     it.mv_y = pu.mv_y;
                     items.push_back(it);
                 });
    }
    return items;
}

// BCW_WEIGHTS (vvc/inter.rs:128): the L1 weight of bcw_idx, clamped to 4 as
// weighted_average_bi does.
constexpr int8_t kBcwWeights[5] = {-2, 3, 4, 5, 10};

// The refinements predict_inter_pu applies to a bi PU (vvc/inter.rs:1763-1800):
// DMVR with default weights on blocks of at least 8x8 and 128 samples, BDOF
// with default weights; the flags carry the caller's enables and POC tests.
bool dmvr_applies(const rmh_bi_pu_desc& p) {
    const int w = 1 << (p.log2_w & 15), h = 1 << (p.log2_h & 15);
    return (p.refine & RMH_BI_DMVR) && p.bcw_idx == 2 && w >= 8 && h >= 8 && w * h >= 128;
}
bool bdof_applies(const rmh_bi_pu_desc& p) { return (p.refine & RMH_BI_BDOF) && p.bcw_idx == 2; }
bool refined(const rmh_bi_pu_desc& p) { return dmvr_applies(p) || bdof_applies(p); }

std::vector<rmh::McBiItem> mc_bi_items(const rmh_mc_bi_frame* frames, const rmh_bi_pu_desc* pus, size_t n_pu) {
    std::vector<rmh::McBiItem> items;
    for (size_t i = 0; i < n_pu; i++) {
        const rmh_bi_pu_desc& pu = pus[i];
        if (refined(pu)) continue;
        pu_tiles(frames[pu.frame].dst, pu.x, pu.y, 1 << (pu.log2_w & 15), 1 << (pu.log2_h & 15),
                 [&](int x0, int y0, int tw, int th, int cw, int ch) {
                     rmh::McBiItem it{};
                     it.x0 = (uint16_t)x0;
                     it.y0 = (uint16_t)y0;
                     it.tw = (uint8_t)tw;
                     it.th = (uint8_t)th;
                     it.cw = (uint8_t)cw;
                     it.ch = (uint8_t)ch;
                     it.frame = pu.frame;
                     it.flags = (pu.log2_w & RMH_PU_ALT_HPEL) ? 1 : 0;
                     it.w1 = kBcwWeights[pu.bcw_idx < 4 ? pu.bcw_idx : 4];
                     it.mv0_x = pu.mv0_x;
                     it.mv0_y = pu.mv0_y;
                     it.mv1_x = pu.mv1_x;
                     it.mv1_y = pu.mv1_y;
                     items.push_back(it);
                 });
    }
    return items;
}

// The work of the refined bi PUs: DMVR jobs, prediction tiles and the
// initial MV pairs (one slot per refined PU).
struct BiRefineWork {
    std::vector<rmh::DmvrJob> jobs;
    std::vector<rmh::McRefItem> items;
    std::vector<int32_t> mvs;
    uint32_t lds = 0;  // dynamic LDS of the DMVR launch
};

BiRefineWork bi_refine_work(const rmh_mc_bi_frame* frames, const rmh_bi_pu_desc* pus, size_t n_pu) {
    BiRefineWork wk;
    for (size_t i = 0; i < n_pu; i++) {
        const rmh_bi_pu_desc& pu = pus[i];
        if (!refined(pu)) continue;
        const uint32_t slot = (uint32_t)(wk.mvs.size() / 4);
        const int lw = pu.log2_w & 15, lh = pu.log2_h & 15, w = 1 << lw, h = 1 << lh;
        wk.mvs.insert(wk.mvs.end(), {pu.mv0_x, pu.mv0_y, pu.mv1_x, pu.mv1_y});
        if (dmvr_applies(pu)) {
            rmh::DmvrJob j{};
            j.x = pu.x;
            j.y = pu.y;
            j.lw = (uint8_t)lw;
            j.lh = (uint8_t)lh;
            j.frame = pu.frame;
            j.slot = slot;
            j.mv[0] = pu.mv0_x;
            j.mv[1] = pu.mv0_y;
            j.mv[2] = pu.mv1_x;
            j.mv[3] = pu.mv1_y;
            wk.jobs.push_back(j);
            wk.lds = std::max(wk.lds, (uint32_t)(2 * (w + 4) * (h + 4)));
        }
        const rmh_plane* dst = frames[pu.frame].dst;
        for (int ty = 0; ty < (h + 15) / 16; ty++)
            for (int tx = 0; tx < (w + 15) / 16; tx++) {
                const int x0 = pu.x + 16 * tx, y0 = pu.y + 16 * ty;
                const bool luma = x0 < (int)dst[0].width && y0 < (int)dst[0].height;
                const bool chroma = x0 / 2 < (int)dst[1].width && y0 / 2 < (int)dst[1].height;
                if (!luma && !chroma) continue;
                rmh::McRefItem it{};
                it.px = pu.x;
                it.py = pu.y;
                it.lw = (uint8_t)lw;
                it.lh = (uint8_t)lh;
                it.tx = (uint8_t)tx;
                it.ty = (uint8_t)ty;
                it.frame = pu.frame;
                it.flags = (uint8_t)(((pu.log2_w & RMH_PU_ALT_HPEL) ? 1 : 0) | (bdof_applies(pu) ? 2 : 0));
                it.w1 = kBcwWeights[pu.bcw_idx < 4 ? pu.bcw_idx : 4];
                it.slot = slot;
                wk.items.push_back(it);
            }
    }
    return wk;
}

uint64_t mc_bytes(const rmh_mc_frame* frames, const rmh_pu_desc* pus, size_t n_pu) {
    uint64_t b = 0;
    for (size_t i = 0; i < n_pu; i++) {
        uint64_t w = 1u << (pus[i].log2_w & 15), h = 1u << (pus[i].log2_h & 15);
        uint64_t bps = frames[pus[i].frame].ref[0].bytes_per_sample;
        b += 2 * bps * (w * h + 2 * (w / 2) * (h / 2)) + 16;
    }
    return b;
}

// Orders work items so that those on 8-bit frames come first (one kernel
// instantiation per sample size); returns how many there are.
template <typename Item, typename Frame>
size_t order_by_sample_size(std::vector<Item>& items, const Frame* frames) {
    auto it = std::stable_partition(items.begin(), items.end(),
                                    [&](const Item& x) { return frames[x.frame].dst[0].bytes_per_sample == 1; });
    return (size_t)(it - items.begin());
}

// Algorithmic bytes of a bi plan: two reference reads and one write per
// sample, plus the 28-byte descriptor.
uint64_t mc_bi_bytes(const rmh_mc_bi_frame* frames, const rmh_bi_pu_desc* pus, size_t n_pu) {
    uint64_t b = 0;
    for (size_t i = 0; i < n_pu; i++) {
        uint64_t w = 1u << (pus[i].log2_w & 15), h = 1u << (pus[i].log2_h & 15);
        uint64_t bps = frames[pus[i].frame].ref0[0].bytes_per_sample;
        b += 3 * bps * (w * h + 2 * (w / 2) * (h / 2)) + sizeof(rmh_bi_pu_desc);
    }
    return b;
}

uint64_t plane_bytes(const rmh_plane& p) { return (uint64_t)p.width * p.height * p.bytes_per_sample; }

// Shadow planes for deblocking: same size and sample type, rows padded to 16
// bytes. shadow_bytes sums what shadow_layout places from `base`.
uint32_t shadow_stride(const rmh_plane& p) { return (uint32_t)(((uint64_t)p.width * p.bytes_per_sample + 15) / 16 * 16 / p.bytes_per_sample); }
size_t shadow_bytes(const rmh_dbk_frame* frames, size_t n_frames) {
    size_t b = 0;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) b += aligned((size_t)shadow_stride(frames[f].planes[c]) * frames[f].planes[c].height * frames[f].planes[c].bytes_per_sample);
    return b;
}
std::vector<rmh_plane> shadow_layout(const rmh_dbk_frame* frames, size_t n_frames, char* base) {
    std::vector<rmh_plane> out;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            rmh_plane q = frames[f].planes[c];
            q.stride = shadow_stride(q);
            q.data = base;
            base += aligned((size_t)q.stride * q.height * q.bytes_per_sample);
            out.push_back(q);
        }
    return out;
}

// Tile numbering of the tiled deblock launch: the frames in an order with
// the 8-bit ones first (stable); words = the running sum of tiles in that
// order (n + 1 entries) followed by the frame at each position (n entries);
// tiles8 = the tiles on 8-bit frames.
struct TileTable {
    std::vector<uint32_t> words;
    uint32_t total = 0, tiles8 = 0;
};
TileTable tile_table(const rmh_dbk_frame* frames, size_t n_frames) {
    std::vector<uint32_t> order(n_frames);
    for (size_t f = 0; f < n_frames; f++) order[f] = (uint32_t)f;
    std::stable_partition(order.begin(), order.end(), [&](uint32_t f) { return frames[f].planes[0].bytes_per_sample == 1; });
    TileTable t;
    t.words.assign(2 * n_frames + 1, 0);
    for (size_t k = 0; k < n_frames; k++) {
        const rmh_plane& y = frames[order[k]].planes[0];
        t.words[k + 1] = t.words[k] + rmh::deblock_tiles(y.width, y.height);
        t.words[n_frames + 1 + k] = order[k];
        if (y.bytes_per_sample == 1) t.tiles8 = t.words[k + 1];
    }
    t.total = t.words[n_frames];
    return t;
}


// TUs the tiled kernel transforms: 2-D with both sides >= 4, no transform
// skip, 8-byte aligned block. Those with LFNST or dependent quantisation that
// the tiles cannot apply themselves (tile_lfnst, tile_dq) are dequantised (and
// LFNST-ed) first by the prepass, one workgroup per TU, and the tiled kernel
// reads its output.
bool tile_eligible(const rmh_tu_desc& t) {
    return t.log2_w >= 2 && t.log2_h >= 2 && !(t.flags & RMH_TU_TRANSFORM_SKIP) && (t.offset & 3) == 0 && !t.joint;
}
bool needs_prepass(const rmh_tu_desc& t);

// LFNST TUs whose coded region covers the 4x4 LFNST input (and, under
// dependent quantisation, is tile_dq): the tiled kernel applies the inverse
// LFNST itself, in LDS, in tiles of their own.
bool tile_lfnst(const rmh_tu_desc& t);

// Dependent quantisation in the tiled kernel: the TU's coded region is its
// whole scan region (min(w, 32) x min(h, 32)), so the levels the automaton
// runs over are exactly the ones the tile loads.
bool tile_dq(const rmh_tu_desc& t);

bool needs_prepass(const rmh_tu_desc& t) {
    const bool lfnst_ok = t.lfnst_idx == 0 || tile_lfnst(t);
    const bool dq_ok = !(t.flags & RMH_TU_DEP_QUANT) || tile_dq(t);
    return !(lfnst_ok && dq_ok);
}

// The LFNST kernel matrix (set * 2 + lfnst_idx - 1, set from the intra mode
// as lfnst_set, vvc/lfnst.rs:85-129) in bits 0-2, the transposition flag in
// bit 3 and bit 4 set (the TU applies LFNST), as the tiled kernel reads them.
uint8_t lfnst_code(const rmh_tu_desc& d) {
    const int m = d.lfnst_mode;
    const int set = m <= 1 ? 0 : m <= 12 ? 1 : m <= 23 ? 2 : m <= 44 ? 3 : m <= 55 ? 2 : 1;
    const bool tp = (m >= 35 && m < 67) || m >= 81;
    return (uint8_t)((set * 2 + (d.lfnst_idx - 1)) | (tp ? 8 : 0) | 16);
}

// Coded region of a TU after the default, MTS, caller and LFNST zero-outs
// (vvc/dct.rs:1063-1077, decode_general.rs:4191-4236), as the generic kernel
// derives it.
void coded_region(const rmh_tu_desc& d, int* cw, int* ch) {
    const int W = 1 << d.log2_w, H = 1 << d.log2_h;
    int sw = (d.tr_hor != RMH_TR_DCT2 && W == 32) ? 16 : (W > 32 ? W - 32 : 0);
    int sh = (d.tr_ver != RMH_TR_DCT2 && H == 32) ? 16 : (H > 32 ? H - 32 : 0);
    sw = std::max<int>(sw, d.skip_w);
    sh = std::max<int>(sh, d.skip_h);
    if (d.lfnst_idx) {
        const int nz = (W == 4 || H == 4) ? 4 : 8;
        sw = std::max(sw, W - std::min(nz, W));
        sh = std::max(sh, H - std::min(nz, H));
    }
    *cw = W - sw;
    *ch = H - sh;
}

bool tile_lfnst(const rmh_tu_desc& t) {
    if (t.lfnst_idx == 0 || ((t.flags & RMH_TU_DEP_QUANT) && !tile_dq(t))) return false;
    int cw = 0, ch = 0;
    coded_region(t, &cw, &ch);
    return cw >= 4 && ch >= 4;
}

bool tile_dq(const rmh_tu_desc& t) {
    if (!(t.flags & RMH_TU_DEP_QUANT)) return false;
    int cw = 0, ch = 0;
    coded_region(t, &cw, &ch);
    return cw == std::min(1 << t.log2_w, 32) && ch == std::min(1 << t.log2_h, 32);
}

// The prepass list of a residual batch: TU indices (in the order of `order`
// after n_tu) and each one's element offset in the prepass buffer.
struct Prepass {
    std::vector<uint32_t> off;      // per prepass TU
    std::vector<uint32_t> off_of;   // per TU index: its prepass offset, or UINT32_MAX
    size_t elements = 0;            // prepass buffer size, int16 elements
};

// Largest autotuned split: a class's tiles hold residual_tile_tus >> split TUs.
constexpr int kMaxTileSplit = 3;

// Tiles of the fast TUs (largest class first) from the class ranges of
// `order`, each class split as `split` says. The last lfnst[lw][lh] TUs of a
// class range apply LFNST or dependent quantisation in the kernel: they get
// tool tiles of their own, marked with rmh::kTileLfnst.
// Plain tiles come first (their number is returned), then the tool tiles:
// each kind is one launch. A tile also ends before every position of `breaks`
// (ascending positions in `order`), so that it holds one transform pair
// where the kernel needs that.
// Recon plans (`recon`) have no uniform tiles.
uint32_t tiles_from_classes(const uint32_t (*first)[7], const uint32_t (*count)[7], const uint32_t (*lfnst)[7],
                            const uint8_t (*split)[7], const std::vector<uint32_t>& breaks, bool recon,
                            std::vector<rmh::TileItem>& tiles, uint32_t* lds) {
    tiles.clear();
    *lds = 0;
    uint32_t n_plain = 0;
    for (int part = 0; part < 2; part++) {
        for (int area = 12; area >= 4; area--)
            for (int lw = 2; lw <= 6; lw++) {
                const int lh = area - lw;
                if (lh < 2 || lh > 6 || count[lw][lh] == 0) continue;
                const bool uni = !part && !recon && rmh::residual_uniform_class(lw, lh);
                const int per = std::max(1, rmh::residual_tile_tus(lw, lh, uni) >> split[lw][lh]);
                const uint32_t plain = count[lw][lh] - lfnst[lw][lh];
                const uint32_t a = part ? plain : 0, b = part ? count[lw][lh] : plain;
                for (uint32_t f = a; f < b;) {
                    uint32_t cnt = std::min<uint32_t>(per, b - f);
                    const uint32_t pos = first[lw][lh] + f;
                    auto nb = std::upper_bound(breaks.begin(), breaks.end(), pos);
                    if (nb != breaks.end() && *nb < pos + cnt) cnt = *nb - pos;
                    tiles.push_back({pos, (uint16_t)cnt, (uint8_t)(lw | (part ? rmh::kTileLfnst : 0)), (uint8_t)lh});
                    f += cnt;
                }
                if (b > a) *lds = std::max(*lds, rmh::residual_tile_lds(lw, lh, uni));
            }
        if (part == 0) n_plain = (uint32_t)tiles.size();
    }
    return n_plain;
}

// Groups tile-eligible TUs by size class (largest class first), appends the
// remaining TUs and then the prepass TUs (so order holds n_tu + prepass
// entries); fills the class ranges, the unsplit tile list, the prepass
// offsets and the tile breaks (the plain TUs of a uniform class are ordered
// by transform pair, and each new pair starts a tile).
uint32_t build_residual_tiles(const rmh_tu_desc* tus, size_t n_tu, std::vector<uint32_t>& order,
                          std::vector<rmh::TileItem>& tiles, size_t* n_fast, uint32_t* lds,
                          uint32_t (*first)[7], uint32_t (*count)[7], uint32_t (*lfnst)[7], Prepass* prep,
                          std::vector<uint32_t>* breaks, bool recon) {
    std::vector<uint32_t> cls[7][7], cls_l[7][7];
    std::vector<uint32_t> rest;
    prep->off.clear();
    prep->off_of.assign(n_tu, UINT32_MAX);
    prep->elements = 0;
    std::vector<uint32_t> prep_list;
    for (size_t i = 0; i < n_tu; i++) {
        if (tile_eligible(tus[i])) {
            (tile_lfnst(tus[i]) || tile_dq(tus[i]) ? cls_l : cls)[tus[i].log2_w][tus[i].log2_h].push_back((uint32_t)i);
            if (needs_prepass(tus[i])) {
                prep_list.push_back((uint32_t)i);
                prep->off.push_back((uint32_t)prep->elements);
                prep->off_of[i] = (uint32_t)prep->elements;
                prep->elements += (size_t)1 << (tus[i].log2_w + tus[i].log2_h);  // a multiple of 16
            }
        } else {
            rest.push_back((uint32_t)i);
        }
    }
    order.clear();
    breaks->clear();
    for (int lw = 0; lw < 7; lw++)
        for (int lh = 0; lh < 7; lh++) first[lw][lh] = count[lw][lh] = lfnst[lw][lh] = 0;
    auto pair_of = [&](uint32_t i) { return tus[i].tr_hor * 3 + tus[i].tr_ver; };
    for (int area = 12; area >= 4; area--)
        for (int lw = 2; lw <= 6; lw++) {
            int lh = area - lw;
            if (lh < 2 || lh > 6) continue;
            // plain tiles of a uniform class run the uniform-matrix passes,
            // which need every TU of a tile to share its transform pair
            if (rmh::residual_uniform_class(lw, lh)) {
                std::vector<uint32_t>& c = cls[lw][lh];
                std::stable_sort(c.begin(), c.end(), [&](uint32_t x, uint32_t y) { return pair_of(x) < pair_of(y); });
                for (size_t k = 1; k < c.size(); k++)
                    if (pair_of(c[k]) != pair_of(c[k - 1])) breaks->push_back((uint32_t)(order.size() + k));
            }
            first[lw][lh] = (uint32_t)order.size();
            count[lw][lh] = (uint32_t)(cls[lw][lh].size() + cls_l[lw][lh].size());
            lfnst[lw][lh] = (uint32_t)cls_l[lw][lh].size();
            order.insert(order.end(), cls[lw][lh].begin(), cls[lw][lh].end());
            order.insert(order.end(), cls_l[lw][lh].begin(), cls_l[lw][lh].end());
        }
    *n_fast = order.size();
    order.insert(order.end(), rest.begin(), rest.end());
    order.insert(order.end(), prep_list.begin(), prep_list.end());
    const uint8_t none[7][7] = {};
    return tiles_from_classes(first, count, lfnst, none, *breaks, recon, tiles, lds);
}

// Tiles a plan can need at the largest split (each class's LFNST TUs start
// a tile of their own).
size_t tile_capacity(const uint32_t (*count)[7]) {
    size_t n = 0;
    for (int lw = 2; lw <= 6; lw++)
        for (int lh = 2; lh <= 6; lh++) {
            const size_t per = std::max(1, rmh::residual_tile_tus(lw, lh) >> kMaxTileSplit);
            n += (count[lw][lh] + per - 1) / per + 1 + (rmh::residual_uniform_class(lw, lh) ? 8 : 0);  // + transform-pair breaks
        }
    return n;
}

// Dequantisation scales of dequantize_rect_bd (vvc/dct.rs:37-40), plain and
// for sqrt(2)-scaled (odd log2 w + log2 h) blocks.
const uint8_t kInvQuantScale[2][6] = {{40, 45, 51, 57, 64, 72}, {57, 64, 72, 80, 90, 102}};

// Tiled-kernel parameters of order[0 .. n_fast) (the fast TUs); prepass TUs
// read their prepass block with unit scale.
std::vector<rmh::TileTu> build_tile_tus(const rmh::DeviceTables& tabs, const rmh_tu_desc* tus,
                                        const std::vector<uint32_t>& order, size_t n_fast, const Prepass& prep) {
    std::vector<rmh::TileTu> out(n_fast);
    for (size_t i = 0; i < n_fast; i++) {
        const rmh_tu_desc& d = tus[order[i]];
        const int lw = d.log2_w, lh = d.log2_h;
        int cw = 0, ch = 0;
        coded_region(d, &cw, &ch);
        const int qp = d.qp < 4 ? 4 : d.qp;
        const int lsum = lw + lh, sqrt2 = lsum & 1;
        const int tr_shift = 15 - d.bit_depth - (lsum >> 1) - sqrt2;
        const int rshift = 6 - (tr_shift + qp / 6);
        rmh::TileTu t{};
        t.off = d.offset;
        t.src = d.offset;
        t.cw = (uint8_t)cw;
        t.ch = (uint8_t)ch;
        t.scale = kInvQuantScale[sqrt2][qp % 6];
        t.shifts = (uint8_t)((rshift < 0 ? -rshift : 0) | ((rshift > 0 ? rshift : 0) << 4));
        if (tile_dq(d)) {
            // dequantize_dep_quant_bd (vvc/dct.rs:1191-1248): the scale of
            // QP + 1 and the shift 7 - (QP + 1) / 6 - tr_shift
            const int qdq = qp + 1, shift = 7 - qdq / 6 - tr_shift;
            t.scale = kInvQuantScale[sqrt2][qdq % 6];
            t.shifts = (uint8_t)((shift < 0 ? -shift : 0) | ((shift > 0 ? shift : 0) << 4));
            t.dq = 1;
        }
        if (tile_lfnst(d)) t.lfnst = lfnst_code(d);
        if (prep.off_of[order[i]] != UINT32_MAX) {
            t.src = prep.off_of[order[i]];
            t.scale = 1;
            t.shifts = 0;
            t.prepped = 1;
        }
        t.mh = (uint16_t)tabs.col16_off[d.tr_hor][lw];
        t.mv = (uint16_t)tabs.col16_off[d.tr_ver][lh];
        t.shift2 = (uint8_t)(20 - d.bit_depth);
        out[i] = t;
    }
    return out;
}

// Recon targets of order[0 .. n_fast) (device picture addresses).
std::vector<rmh::TileDst> build_tile_dst(const std::vector<uint32_t>& order, size_t n_fast, const rmh_tu_place* place,
                                         const rmh_recon_frame* frames) {
    std::vector<rmh::TileDst> out(n_fast);
    for (size_t i = 0; i < n_fast; i++) {
        const rmh_tu_place& pl = place[order[i]];
        const rmh_plane& p = frames[pl.frame].planes[pl.plane];
        rmh::TileDst t{};
        t.base = (char*)p.data + ((size_t)pl.y * p.stride + pl.x) * p.bytes_per_sample;
        t.stride = (int32_t)p.stride;
        t.bps = p.bytes_per_sample;
        t.maxv = (uint16_t)((1 << p.bit_depth) - 1);
        out[i] = t;
    }
    return out;
}

// Parts of a whole residual-plan run (a frame-step graph runs them on
// branches of their own): the dequantising prepass, the tiled launches and
// the generic kernel for the remaining TUs.
enum { kPartPrepass = 1, kPartTiles = 2, kPartGeneric = 4, kPartsAll = 7 };

// True when a run of the whole plan on these buffers takes the tiled path.
bool residual_tiled_run(const rmh_plan* plan, const int16_t* d_coeff, const int16_t* d_residual) {
    const bool aligned8 = ((uintptr_t)d_coeff & 7) == 0 && (plan->d_place || ((uintptr_t)d_residual & 7) == 0);
    return plan->kind == kResidual && plan->d_order && aligned8;
}

rmh_status run_plan_locked_free(rmh_plan* plan, const int16_t* d_coeff, int16_t* d_residual, size_t first, size_t count,
                                hipStream_t s, std::string* err, int parts = kPartsAll) {
    hipError_t e = hipSuccess;
    if (plan->kind == kResidual) {
        rmh::ResidualArgs a{};
        a.tus = plan->d_tus;
        a.coeff = d_coeff;
        a.residual = d_residual;
        a.mat = plan->ctx->tabs.mat;
        a.lfnst8 = plan->ctx->tabs.lfnst8;
        a.lfnst4 = plan->ctx->tabs.lfnst4;
        a.scans = plan->ctx->tabs.scans;
        memcpy(a.mat_off, plan->ctx->tabs.mat_off, sizeof a.mat_off);
        memcpy(a.scan_off, plan->ctx->tabs.scan_off, sizeof a.scan_off);
        memcpy(a.scan_len, plan->ctx->tabs.scan_len, sizeof a.scan_len);
        a.col16 = plan->ctx->tabs.col16;
        memcpy(a.col16_off, plan->ctx->tabs.col16_off, sizeof a.col16_off);
        memcpy(a.eo_off, plan->ctx->tabs.eo_off, sizeof a.eo_off);
        a.place = plan->d_place;
        a.frames = plan->d_rframes;
        a.tile_tus = plan->d_tile_tus;
        a.tile_dst = plan->d_tile_dst;
        const bool whole = first == 0 && count == plan->n_tu;
        if (whole && residual_tiled_run(plan, d_coeff, d_residual)) {
            // the prepass, the tiled kernel for the 2-D TUs, the general kernel for the rest
            a.order = plan->d_order;
            if (plan->n_prep && (parts & kPartPrepass)) {
                a.prep_out = plan->d_prep;
                a.prep_off = plan->d_prep_off;
                a.first_tu = (uint32_t)plan->n_tu;
                a.n_tu = (uint32_t)plan->n_prep;
                e = rmh::launch_residual(a, s);
                a.prep_out = nullptr;
                a.prep_off = nullptr;
            }
            a.prep = plan->d_prep;
            if (e == hipSuccess && (parts & kPartTiles))
                e = rmh::launch_residual_tiles(a, plan->d_tiles, (uint32_t)plan->n_tiles, plan->n_tiles_plain,
                                               plan->tile_lds, s);
            if (e == hipSuccess && plan->n_tu > plan->n_fast && (parts & kPartGeneric)) {
                a.first_tu = (uint32_t)plan->n_fast;
                a.n_tu = (uint32_t)(plan->n_tu - plan->n_fast);
                e = rmh::launch_residual(a, s);
            }
        } else {
            a.order = nullptr;
            a.first_tu = (uint32_t)first;
            a.n_tu = (uint32_t)count;
            e = rmh::launch_residual(a, s);
        }
    } else if (plan->kind == kMc) {
        if (plan->d_bi_frames) {
            // DMVR writes the refined pairs that the refined PUs' prediction reads
            e = rmh::launch_dmvr(plan->d_bi_frames, plan->d_dmvr, (uint32_t)plan->n_dmvr, plan->dmvr_lds, plan->d_mvs, s);
            if (e == hipSuccess)
                e = rmh::launch_mc_bi_refined(plan->d_bi_frames, plan->d_ref_items, (uint32_t)plan->n_ref_items,
                                              plan->d_mvs, s);
            if (e == hipSuccess)
                e = rmh::launch_mc_bi(plan->d_bi_frames, plan->d_bi_items, (uint32_t)plan->n_items,
                                      (uint32_t)plan->n_items8, s, plan->mc_waves);
        } else {
            e = rmh::launch_mc(plan->d_frames, plan->d_items, (uint32_t)plan->n_items, (uint32_t)plan->n_items8, s,
                               plan->mc_waves);
        }
    } else if (plan->kind == kSao) {
        e = rmh::launch_sao(plan->d_sao_jobs, plan->d_sao_block_job, plan->sao_range, s);
    } else if (plan->kind == kEdge) {
        e = rmh::launch_edge_map(plan->d_edge_frames, plan->d_cell_start, (uint32_t)plan->n_edge, plan->total_cells, s);
    } else if (plan->kind == kForward) {
        rmh::FwdArgs a{};
        a.tus = plan->d_tus;
        a.residual = d_coeff;
        a.levels = d_residual;
        a.mat = plan->ctx->tabs.mat;
        a.lfnst8 = plan->ctx->tabs.lfnst8;
        a.lfnst4 = plan->ctx->tabs.lfnst4;
        a.scans = plan->ctx->tabs.scans;
        memcpy(a.mat_off, plan->ctx->tabs.mat_off, sizeof a.mat_off);
        memcpy(a.scan_off, plan->ctx->tabs.scan_off, sizeof a.scan_off);
        memcpy(a.scan_len, plan->ctx->tabs.scan_len, sizeof a.scan_len);
        a.n_tu = (uint32_t)plan->n_tu;
        e = rmh::launch_forward(a, s);
    } else if (plan->kind == kGpm) {
        e = rmh::launch_gpm(plan->d_gpm_frames, plan->d_gpm_items, (uint32_t)plan->n_gpm_items, plan->gpm_scale[0],
                            plan->gpm_scale[1], s);
    } else if (plan->kind == kLmcs) {
        e = rmh::launch_lmcs(plan->d_lmcs_jobs, plan->d_lmcs_block_job, plan->lmcs_range, s);
    } else if (plan->kind == kAlf) {
        e = rmh::launch_alf(plan->d_alf_luma, plan->d_alf_luma_bj, plan->alf_luma_range, plan->d_alf_chroma,
                            plan->d_alf_chroma_bj, plan->alf_chroma_range, s);
    } else {
        e = rmh::launch_deblock(plan->d_dbk, plan->d_shadows, plan->d_tile_start, plan->d_frame_of, (uint32_t)plan->n_dbk,
                                plan->total_tiles, plan->tiles8, plan->max_rows, plan->dbk_copy, s);
    }
    if (e != hipSuccess) {
        *err = std::string("kernel launch failed: ") + hipGetErrorString(e);
        return RMH_EXECUTION;
    }
    return RMH_OK;
}

rmh_status run_plan(rmh_plan* plan, const int16_t* d_coeff, int16_t* d_residual, size_t first, size_t count, void* stream) {
    rmh_ctx* ctx = plan->ctx;
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    std::string err;
    rmh_status s = run_plan_locked_free(plan, d_coeff, d_residual, first, count, (hipStream_t)stream, &err);
    if (s != RMH_OK) return fail(ctx, s, err);
    return RMH_OK;
}

}  // namespace

// ======================================================================
// C ABI
// ======================================================================

RMH_EXPORT rmh_status rmh_create(const rmh_config* cfg, rmh_ctx** out) {
    if (!out) return RMH_INVALID_INPUT;
    *out = nullptr;
    rmh_config c{0, 0, 0};
    if (cfg) c = *cfg;
    if (c.flags != 0) return RMH_INVALID_INPUT;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RMH_UNAVAILABLE;
    if (c.device < 0 || c.device >= n) return RMH_INVALID_INPUT;
    rmh_ctx* ctx = new rmh_ctx();
    ctx->device = c.device;
    ctx->timeout_ms = c.timeout_ms ? c.timeout_ms : 30000;
    ctx->host.pinned = true;
    DeviceGuard dg(ctx->device);
    if (hipGetDeviceProperties(&ctx->props, ctx->device) != hipSuccess ||
        strncmp(ctx->props.gcnArchName, "gfx950", 6) != 0) {
        delete ctx;
        return RMH_UNAVAILABLE;
    }
    ctx->scratch = scratch_address(ctx->device);
    if (!ctx->scratch || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        build_tables(ctx) != hipSuccess) {
        if (ctx->stream) hipStreamDestroy(ctx->stream);
        if (ctx->d_tables) hipFree(ctx->d_tables);
        delete ctx;
        return RMH_UNAVAILABLE;
    }
    *out = ctx;
    return RMH_OK;
}

RMH_EXPORT void rmh_destroy(rmh_ctx* ctx) {
    if (!ctx) return;
    DeviceGuard dg(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ctx->dev.release();
    ctx->host.release();
    if (ctx->d_tables) hipFree(ctx->d_tables);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

RMH_EXPORT const char* rmh_last_error(const rmh_ctx* ctx) {
    static thread_local std::string copy;
    if (!ctx) return "null context";
    rmh_ctx* c = const_cast<rmh_ctx*>(ctx);
    std::lock_guard<std::mutex> lk(c->mu);
    copy = c->last_error;
    return copy.c_str();
}

RMH_EXPORT rmh_status rmh_info(rmh_ctx* ctx, rmh_device_info* out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    memset(out, 0, sizeof *out);
    snprintf(out->name, sizeof out->name, "%s", ctx->props.name);
    snprintf(out->arch, sizeof out->arch, "%s", ctx->props.gcnArchName);
    out->compute_units = ctx->props.multiProcessorCount;
    out->max_clock_khz = ctx->props.clockRate;
    out->global_mem_bytes = ctx->props.totalGlobalMem;
    out->abi_version = RMH_ABI_VERSION;
    return RMH_OK;
}

RMH_EXPORT void rmh_disable(rmh_ctx* ctx) {
    if (ctx) ctx->disabled.store(true);
}

RMH_EXPORT int rmh_is_disabled(const rmh_ctx* ctx) { return ctx ? (int)ctx->disabled.load() : 1; }

RMH_EXPORT void rmh_inject_failures(rmh_ctx* ctx, uint32_t count) {
    if (ctx) ctx->inject.store(count);
}

RMH_EXPORT rmh_status rmh_transform_matrix(int32_t type, int32_t n, int32_t* out) {
    if (!out || n < 2 || n > 64 || (n & (n - 1)) || !matrix_ok(type, n)) return RMH_INVALID_INPUT;
    bool ok;
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) out[k * n + j] = rmh::transform_entry(type, n, k, j, &ok);
    return RMH_OK;
}

// ---------------------------------------------------------------- residual

namespace {

rmh_status create_residual_plan(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu, size_t n_coeff, const rmh_tu_place* place,
                                const rmh_recon_frame* frames, size_t n_frames, rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    rmh_status v = validate_tus(ctx, tus, n_tu, n_coeff, false);
    if (v != RMH_OK) return v;
    const bool recon = place != nullptr || frames != nullptr;
    if (recon && (v = validate_place(ctx, tus, n_tu, place, frames, n_frames)) != RMH_OK) return v;
    if (!recon)
        for (size_t i = 0; i < n_tu; i++)
            if (tus[i].joint)
                return fail(ctx, RMH_INVALID_INPUT, "TU " + std::to_string(i) + ": joint Cb-Cr residuals need a recon plan");
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kResidual;
    p->n_tu = n_tu;
    p->n_coeff = n_coeff;
    p->bytes = residual_bytes(tus, n_tu);
    if (recon) {
        // the recon add reads and writes each sample instead of writing an int16 residual
        uint64_t extra = 0;
        for (size_t i = 0; i < n_tu; i++)
            extra += ((uint64_t)frames[place[i].frame].planes[place[i].plane].bytes_per_sample * 2 - 2) << (tus[i].log2_w + tus[i].log2_h);
        p->bytes += extra + 8 * (uint64_t)n_tu;
    }
    if (n_tu) {
        std::vector<uint32_t> order;
        std::vector<rmh::TileItem> tiles;
        Prepass prep;
        p->n_tiles_plain = build_residual_tiles(tus, n_tu, order, tiles, &p->n_fast, &p->tile_lds, p->cls_first,
                                                p->cls_count, p->cls_lfnst, &prep, &p->tile_breaks, recon);
        p->n_prep = prep.off.size();
        p->tile_capacity = std::max(tiles.size(), tile_capacity(p->cls_count));
        size_t tb = aligned(n_tu * sizeof(rmh_tu_desc)), ob = aligned(order.size() * 4),
               ib = aligned(p->tile_capacity * sizeof(rmh::TileItem));
        size_t pb = recon ? aligned(n_tu * sizeof(rmh_tu_place)) : 0, fb = recon ? aligned(n_frames * sizeof(rmh_recon_frame)) : 0;
        const std::vector<rmh::TileTu> ttu = build_tile_tus(ctx->tabs, tus, order, p->n_fast, prep);
        std::vector<rmh::TileDst> tdst;
        if (recon) tdst = build_tile_dst(order, p->n_fast, place, frames);
        size_t ub = aligned(ttu.size() * sizeof(rmh::TileTu)), db = aligned(tdst.size() * sizeof(rmh::TileDst));
        size_t qb = aligned(prep.off.size() * 4), xb = aligned(prep.elements * sizeof(int16_t));
        hipError_t e = hipMalloc(&p->d_mem, tb + ob + ib + pb + fb + ub + db + qb + xb + 256);
        char* base = (char*)p->d_mem;
        if (e == hipSuccess) e = hipMemcpy(base, tus, n_tu * sizeof(rmh_tu_desc), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(base + tb, order.data(), order.size() * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess && !tiles.empty())
            e = hipMemcpy(base + tb + ob, tiles.data(), tiles.size() * sizeof(rmh::TileItem), hipMemcpyHostToDevice);
        if (e == hipSuccess && recon) e = hipMemcpy(base + tb + ob + ib, place, n_tu * sizeof(rmh_tu_place), hipMemcpyHostToDevice);
        if (e == hipSuccess && recon && n_frames)
            e = hipMemcpy(base + tb + ob + ib + pb, frames, n_frames * sizeof(rmh_recon_frame), hipMemcpyHostToDevice);
        char* ubase = base + tb + ob + ib + pb + fb;
        if (e == hipSuccess && !ttu.empty()) e = hipMemcpy(ubase, ttu.data(), ttu.size() * sizeof(rmh::TileTu), hipMemcpyHostToDevice);
        if (e == hipSuccess && !tdst.empty())
            e = hipMemcpy(ubase + ub, tdst.data(), tdst.size() * sizeof(rmh::TileDst), hipMemcpyHostToDevice);
        if (e == hipSuccess && !prep.off.empty())
            e = hipMemcpy(ubase + ub + db, prep.off.data(), prep.off.size() * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (p->d_mem) hipFree(p->d_mem);
            delete p;
            return fail(ctx, RMH_EXECUTION, std::string("descriptor upload failed: ") + hipGetErrorString(e));
        }
        p->d_tus = (const rmh_tu_desc*)base;
        p->d_order = (const uint32_t*)(base + tb);
        p->d_tiles = (rmh::TileItem*)(base + tb + ob);
        p->n_tiles = tiles.size();
        if (recon) {
            p->d_place = (const rmh_tu_place*)(base + tb + ob + ib);
            p->d_rframes = (const rmh_recon_frame*)(base + tb + ob + ib + pb);
            p->d_tile_dst = (const rmh::TileDst*)(ubase + ub);
        }
        p->d_tile_tus = (const rmh::TileTu*)ubase;
        if (p->n_prep) {
            p->d_prep_off = (const uint32_t*)(ubase + ub + db);
            p->d_prep = (int16_t*)(ubase + ub + db + qb);
        }
    }
    *out = p;
    return RMH_OK;
}

}  // namespace

RMH_EXPORT rmh_status rmh_residual_plan_create(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu, size_t n_coeff,
                                               rmh_plan** out) {
    return create_residual_plan(ctx, tus, n_tu, n_coeff, nullptr, nullptr, 0, out);
}

RMH_EXPORT rmh_status rmh_residual_recon_plan_create(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu, size_t n_coeff,
                                                     const rmh_tu_place* place, const rmh_recon_frame* frames,
                                                     size_t n_frames, rmh_plan** out) {
    if (!place || !frames) {
        if (out) *out = nullptr;
        return ctx ? fail(ctx, RMH_INVALID_INPUT, "recon plan without placement or frames") : RMH_INVALID_INPUT;
    }
    return create_residual_plan(ctx, tus, n_tu, n_coeff, place, frames, n_frames, out);
}

RMH_EXPORT rmh_status rmh_residual_plan_run(rmh_plan* plan, const int16_t* d_coeff, int16_t* d_residual, void* stream) {
    if (!plan || plan->kind != kResidual) return RMH_INVALID_INPUT;
    if (plan->n_tu == 0) return RMH_OK;
    if (!d_coeff || (!d_residual && !plan->d_place)) return fail(plan->ctx, RMH_INVALID_INPUT, "null device buffer");
    return run_plan(plan, d_coeff, d_residual, 0, plan->n_tu, stream);
}

RMH_EXPORT rmh_status rmh_residual_plan_run_range(rmh_plan* plan, size_t first, size_t count, const int16_t* d_coeff,
                                                  int16_t* d_residual, void* stream) {
    if (!plan || plan->kind != kResidual) return RMH_INVALID_INPUT;
    if (first > plan->n_tu || count > plan->n_tu - first) return fail(plan->ctx, RMH_INVALID_INPUT, "TU range outside the plan");
    if (count == 0) return RMH_OK;
    if (!d_coeff || (!d_residual && !plan->d_place)) return fail(plan->ctx, RMH_INVALID_INPUT, "null device buffer");
    return run_plan(plan, d_coeff, d_residual, first, count, stream);
}

RMH_EXPORT rmh_status rmh_residual_plan_autotune(rmh_plan* plan, const int16_t* d_coeff, int16_t* d_residual,
                                                 void* stream, uint32_t reps, uint8_t* split_out) {
    if (!plan || plan->kind != kResidual) return RMH_INVALID_INPUT;
    rmh_ctx* ctx = plan->ctx;
    if (!d_coeff || (!d_residual && !plan->d_place)) return fail(ctx, RMH_INVALID_INPUT, "null device buffer");
    if (((uintptr_t)d_coeff & 7) || (!plan->d_place && ((uintptr_t)d_residual & 7)))
        return fail(ctx, RMH_INVALID_INPUT, "tiled kernel needs 8-byte aligned buffers");
    // a captured graph replays the tile list and grid size it was built with
    if (plan->graph_refs.load() != 0)
        return fail(ctx, RMH_INVALID_INPUT, "plan is captured in a graph: destroy the graph before autotuning");
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    reps = reps ? std::min<uint32_t>(reps, 100) : 5;
    rmh::ResidualArgs a{};
    a.tus = plan->d_tus;
    a.coeff = d_coeff;
    a.residual = d_residual;
    a.mat = ctx->tabs.mat;
    a.lfnst8 = ctx->tabs.lfnst8;
    a.lfnst4 = ctx->tabs.lfnst4;
    a.scans = ctx->tabs.scans;
    memcpy(a.mat_off, ctx->tabs.mat_off, sizeof a.mat_off);
    memcpy(a.scan_off, ctx->tabs.scan_off, sizeof a.scan_off);
    memcpy(a.scan_len, ctx->tabs.scan_len, sizeof a.scan_len);
    a.col16 = ctx->tabs.col16;
    memcpy(a.col16_off, ctx->tabs.col16_off, sizeof a.col16_off);
    memcpy(a.eo_off, ctx->tabs.eo_off, sizeof a.eo_off);
    a.place = plan->d_place;
    a.frames = plan->d_rframes;
    a.tile_tus = plan->d_tile_tus;
    a.tile_dst = plan->d_tile_dst;
    a.order = plan->d_order;
    hipError_t e = hipSuccess;
    if (plan->n_prep) {
        // the prepass output the tiled launches read
        a.prep_out = plan->d_prep;
        a.prep_off = plan->d_prep_off;
        a.first_tu = (uint32_t)plan->n_tu;
        a.n_tu = (uint32_t)plan->n_prep;
        e = rmh::launch_residual(a, s);
        a.prep_out = nullptr;
        a.prep_off = nullptr;
    }
    a.prep = plan->d_prep;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // coordinate descent over the classes (largest first): each candidate split
    // of one class is timed on the whole plan's launch as the fastest of three
    // batches of reps launches (run-to-run noise is about 2 %); a change is
    // kept when it saves at least 2 %
    auto time_plan = [&](const uint8_t (*split)[7], float* ms) -> hipError_t {
        std::vector<rmh::TileItem> tiles;
        uint32_t lds = 0;
        const uint32_t n_plain = tiles_from_classes(plan->cls_first, plan->cls_count, plan->cls_lfnst, split, plan->tile_breaks,
                                                    plan->d_place != nullptr, tiles, &lds);
        hipError_t r = hipMemcpy(plan->d_tiles, tiles.data(), tiles.size() * sizeof(rmh::TileItem), hipMemcpyHostToDevice);
        for (uint32_t k = 0; k < 2 && r == hipSuccess; k++)
            r = rmh::launch_residual_tiles(a, plan->d_tiles, (uint32_t)tiles.size(), n_plain, lds, s);
        float fastest = 0.0f;
        for (int batch = 0; batch < 3 && r == hipSuccess; batch++) {
            r = hipEventRecord(e0, s);
            for (uint32_t k = 0; k < reps && r == hipSuccess; k++)
                r = rmh::launch_residual_tiles(a, plan->d_tiles, (uint32_t)tiles.size(), n_plain, lds, s);
            if (r == hipSuccess) r = hipEventRecord(e1, s);
            if (r == hipSuccess) r = hipEventSynchronize(e1);
            float t = 0.0f;
            if (r == hipSuccess) r = hipEventElapsedTime(&t, e0, e1);
            if (r == hipSuccess && (batch == 0 || t < fastest)) fastest = t;
        }
        *ms = fastest;
        return r;
    };
    uint8_t cur[7][7];
    memcpy(cur, plan->split, sizeof cur);
    float best = 0.0f;
    if (e == hipSuccess) e = time_plan(cur, &best);
    for (int area = 12; area >= 4 && e == hipSuccess; area--)
        for (int lw = 2; lw <= 6 && e == hipSuccess; lw++) {
            const int lh = area - lw;
            if (lh < 2 || lh > 6 || plan->cls_count[lw][lh] == 0) continue;
            const uint8_t keep = cur[lw][lh];
            uint8_t choice = keep;
            for (int sp = 0; sp <= kMaxTileSplit && e == hipSuccess; sp++) {
                const bool uni = plan->d_place == nullptr && rmh::residual_uniform_class(lw, lh);
                if (sp == keep || (rmh::residual_tile_tus(lw, lh, uni) >> sp) < 1) continue;
                cur[lw][lh] = (uint8_t)sp;
                float ms = 0.0f;
                e = time_plan(cur, &ms);
                if (e == hipSuccess && ms < 0.98f * best) {
                    best = ms;
                    choice = (uint8_t)sp;
                }
            }
            cur[lw][lh] = choice;
        }
    if (e == hipSuccess) memcpy(plan->split, cur, sizeof cur);
    // the whole plan's tiles with the chosen splits
    if (e == hipSuccess) {
        std::vector<rmh::TileItem> tiles;
        plan->n_tiles_plain =
            tiles_from_classes(plan->cls_first, plan->cls_count, plan->cls_lfnst, plan->split, plan->tile_breaks,
                               plan->d_place != nullptr, tiles,
                               &plan->tile_lds);
        e = hipMemcpy(plan->d_tiles, tiles.data(), tiles.size() * sizeof(rmh::TileItem), hipMemcpyHostToDevice);
        plan->n_tiles = tiles.size();
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (e != hipSuccess) return fail(ctx, RMH_EXECUTION, std::string("autotune: ") + hipGetErrorString(e));
    if (split_out)
        for (int lw = 0; lw < 7; lw++)
            for (int lh = 0; lh < 7; lh++) split_out[lw * 7 + lh] = plan->split[lw][lh];
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_residual_batch(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu, const int16_t* coeff,
                                         size_t n_coeff, int16_t* residual) {
    if (!ctx) return RMH_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(ctx->mu);
    rmh_status v = validate_tus(ctx, tus, n_tu, n_coeff, true);
    if (v != RMH_OK) return v;
    if (n_tu == 0) return RMH_OK;
    if (!coeff || !residual) return fail_locked(ctx, RMH_INVALID_INPUT, "null coefficient or residual buffer");
    for (size_t i = 0; i < n_tu; i++)
        if (tus[i].joint)
            return fail_locked(ctx, RMH_INVALID_INPUT, "TU " + std::to_string(i) + ": joint Cb-Cr residuals need a recon plan");
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail_locked(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    size_t tb = n_tu * sizeof(rmh_tu_desc), cb = n_coeff * sizeof(int16_t);
    std::vector<uint32_t> order;
    std::vector<rmh::TileItem> tiles;
    rmh_plan tmp;
    Prepass prep;
    tmp.n_tiles_plain = build_residual_tiles(tus, n_tu, order, tiles, &tmp.n_fast, &tmp.tile_lds, tmp.cls_first,
                                             tmp.cls_count, tmp.cls_lfnst, &prep, &tmp.tile_breaks, false);
    hipError_t e = ctx->dev.reserve(aligned(tb) + 2 * aligned(cb) + aligned(order.size() * 4) +
                                    aligned(tiles.size() * sizeof(rmh::TileItem)) +
                                    aligned(tmp.n_fast * sizeof(rmh::TileTu)) + aligned(prep.off.size() * 4) +
                                    aligned(prep.elements * sizeof(int16_t)));
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("device allocation: ") + hipGetErrorString(e));
    e = ctx->host.reserve(aligned(cb));
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("host staging: ") + hipGetErrorString(e));
    rmh_tu_desc* d_tus = (rmh_tu_desc*)ctx->dev.take(tb);
    int16_t* d_coeff = (int16_t*)ctx->dev.take(cb);
    int16_t* d_res = (int16_t*)ctx->dev.take(cb);
    int16_t* h_res = (int16_t*)ctx->host.take(cb);
    hipStream_t s = ctx->stream;
    uint32_t* d_order = (uint32_t*)ctx->dev.take(order.size() * 4);
    rmh::TileItem* d_tiles = (rmh::TileItem*)ctx->dev.take(tiles.size() * sizeof(rmh::TileItem));
    const std::vector<rmh::TileTu> ttu = build_tile_tus(ctx->tabs, tus, order, tmp.n_fast, prep);
    rmh::TileTu* d_ttu = (rmh::TileTu*)ctx->dev.take(ttu.size() * sizeof(rmh::TileTu));
    uint32_t* d_prep_off = (uint32_t*)ctx->dev.take(prep.off.size() * 4);
    int16_t* d_prep = (int16_t*)ctx->dev.take(prep.elements * sizeof(int16_t));
    if ((e = hipMemcpyAsync(d_tus, tus, tb, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(d_coeff, coeff, cb, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (!tiles.empty() &&
         (e = hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(rmh::TileItem), hipMemcpyHostToDevice, s)) != hipSuccess) ||
        (!ttu.empty() &&
         (e = hipMemcpyAsync(d_ttu, ttu.data(), ttu.size() * sizeof(rmh::TileTu), hipMemcpyHostToDevice, s)) != hipSuccess) ||
        (!prep.off.empty() &&
         (e = hipMemcpyAsync(d_prep_off, prep.off.data(), prep.off.size() * 4, hipMemcpyHostToDevice, s)) != hipSuccess))
        return hip_fail_locked(ctx, e, "upload");
    tmp.ctx = ctx;
    tmp.kind = kResidual;
    tmp.d_tus = d_tus;
    tmp.n_tu = n_tu;
    tmp.d_order = d_order;
    tmp.d_tiles = d_tiles;
    tmp.n_tiles = tiles.size();
    tmp.d_tile_tus = d_ttu;
    tmp.n_prep = prep.off.size();
    tmp.d_prep_off = d_prep_off;
    tmp.d_prep = d_prep;
    std::string err;
    rmh_status r = run_plan_locked_free(&tmp, d_coeff, d_res, 0, n_tu, s, &err);
    if (r != RMH_OK) return fail_locked(ctx, r, err);
    if ((e = hipMemcpyAsync(h_res, d_res, cb, hipMemcpyDeviceToHost, s)) != hipSuccess) return hip_fail_locked(ctx, e, "download");
    r = wait_stream(ctx, s);
    if (r != RMH_OK) return r;
    // samples not covered by any TU keep the caller's values
    for (size_t i = 0; i < n_tu; i++) {
        size_t n = (size_t)1 << (tus[i].log2_w + tus[i].log2_h);
        memcpy(residual + tus[i].offset, h_res + tus[i].offset, n * sizeof(int16_t));
    }
    return RMH_OK;
}

// -------------------------------------------------------------------- inter

RMH_EXPORT rmh_status rmh_mc_plan_create(rmh_ctx* ctx, const rmh_mc_frame* frames, size_t n_frames, const rmh_pu_desc* pus,
                                         size_t n_pu, rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rmh_status v = validate_mc(ctx, frames, n_frames, pus, n_pu);
        if (v != RMH_OK) return v;
    }
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    std::vector<rmh::McItem> items = mc_items(frames, pus, n_pu);
    if (items.size() > 0x7fffffffu) return fail(ctx, RMH_LIMIT_EXCEEDED, "too many inter work items");
    const size_t n8 = order_by_sample_size(items, frames);
    size_t fb = n_frames * sizeof(rmh_mc_frame), pb = n_pu * sizeof(rmh_pu_desc), ib = items.size() * sizeof(rmh::McItem);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kMc;
    p->n_items = items.size();
    p->n_items8 = n8;
    p->bytes = mc_bytes(frames, pus, n_pu);
    hipError_t e = hipMalloc(&p->d_mem, aligned(fb) + aligned(pb) + aligned(ib) + 256);
    char* base = (char*)p->d_mem;
    if (e == hipSuccess && fb) e = hipMemcpy(base, frames, fb, hipMemcpyHostToDevice);
    if (e == hipSuccess && pb) e = hipMemcpy(base + aligned(fb), pus, pb, hipMemcpyHostToDevice);
    if (e == hipSuccess && ib) e = hipMemcpy(base + aligned(fb) + aligned(pb), items.data(), ib, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (p->d_mem) hipFree(p->d_mem);
        delete p;
        return fail(ctx, RMH_EXECUTION, std::string("inter plan upload failed: ") + hipGetErrorString(e));
    }
    p->d_frames = (const rmh_mc_frame*)base;
    p->d_pus = (const rmh_pu_desc*)(base + aligned(fb));
    p->d_items = (const rmh::McItem*)(base + aligned(fb) + aligned(pb));
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_mc_plan_run(rmh_plan* plan, void* stream) {
    if (!plan || plan->kind != kMc) return RMH_INVALID_INPUT;
    if (plan->n_items == 0 && plan->n_ref_items == 0) return RMH_OK;
    return run_plan(plan, nullptr, nullptr, 0, 0, stream);
}

RMH_EXPORT rmh_status rmh_mc_plan_autotune(rmh_plan* plan, void* stream, uint32_t* waves_out) {
    if (!plan || plan->kind != kMc) return RMH_INVALID_INPUT;
    rmh_ctx* ctx = plan->ctx;
    if (plan->graph_refs.load() != 0)
        return fail(ctx, RMH_INVALID_INPUT, "plan is captured in a graph: destroy the graph before autotuning");
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // each shape timed on the plan's MC launch as the fastest of three batches
    // of five launches (run-to-run noise is about 2 %), after two untimed ones
    auto launch = [&](int waves) {
        return plan->d_bi_frames ? rmh::launch_mc_bi(plan->d_bi_frames, plan->d_bi_items, (uint32_t)plan->n_items,
                                                     (uint32_t)plan->n_items8, s, waves)
                                 : rmh::launch_mc(plan->d_frames, plan->d_items, (uint32_t)plan->n_items,
                                                  (uint32_t)plan->n_items8, s, waves);
    };
    auto time_shape = [&](int waves, float* ms) -> hipError_t {
        hipError_t r = hipSuccess;
        for (int k = 0; k < 2 && r == hipSuccess; k++) r = launch(waves);
        float fastest = 0.0f;
        for (int batch = 0; batch < 3 && r == hipSuccess; batch++) {
            r = hipEventRecord(e0, s);
            for (int k = 0; k < 5 && r == hipSuccess; k++) r = launch(waves);
            if (r == hipSuccess) r = hipEventRecord(e1, s);
            if (r == hipSuccess) r = hipEventSynchronize(e1);
            float t = 0.0f;
            if (r == hipSuccess) r = hipEventElapsedTime(&t, e0, e1);
            if (r == hipSuccess && (batch == 0 || t < fastest)) fastest = t;
        }
        *ms = fastest;
        return r;
    };
    int choice = plan->mc_waves;
    float best = 0.0f;
    if (e == hipSuccess && plan->n_items) e = time_shape(choice, &best);
    for (int waves : {4, 2, 1}) {
        if (e != hipSuccess || !plan->n_items || waves == plan->mc_waves) continue;
        float ms = 0.0f;
        e = time_shape(waves, &ms);
        if (e == hipSuccess && ms < 0.98f * best) {
            best = ms;
            choice = waves;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (e != hipSuccess) return fail(ctx, RMH_EXECUTION, std::string("autotune: ") + hipGetErrorString(e));
    plan->mc_waves = choice;
    if (waves_out) *waves_out = (uint32_t)choice;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_mc_batch(rmh_ctx* ctx, const rmh_mc_frame* frames, size_t n_frames, const rmh_pu_desc* pus,
                                   size_t n_pu) {
    if (!ctx) return RMH_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(ctx->mu);
    rmh_status v = validate_mc(ctx, frames, n_frames, pus, n_pu);
    if (v != RMH_OK) return v;
    if (n_pu == 0) return RMH_OK;
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail_locked(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    std::vector<rmh::McItem> items = mc_items(frames, pus, n_pu);
    const size_t n8 = order_by_sample_size(items, frames);
    size_t fb = n_frames * sizeof(rmh_mc_frame), pb = n_pu * sizeof(rmh_pu_desc), ib = items.size() * sizeof(rmh::McItem);
    size_t need = aligned(fb) + aligned(pb) + aligned(ib), hneed = 0;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            need += 2 * aligned(plane_bytes(frames[f].ref[c]));
            hneed += aligned(plane_bytes(frames[f].dst[c]));
        }
    hipError_t e = ctx->dev.reserve(need);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("device allocation: ") + hipGetErrorString(e));
    e = ctx->host.reserve(hneed);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("host staging: ") + hipGetErrorString(e));
    hipStream_t s = ctx->stream;
    rmh_mc_frame* d_frames = (rmh_mc_frame*)ctx->dev.take(fb);
    rmh_pu_desc* d_pus = (rmh_pu_desc*)ctx->dev.take(pb);
    rmh::McItem* d_items = (rmh::McItem*)ctx->dev.take(ib);
    std::vector<rmh_mc_frame> dev_frames(frames, frames + n_frames);
    std::vector<void*> staged;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            const rmh_plane& r = frames[f].ref[c];
            void* dr = ctx->dev.take(plane_bytes(r));
            void* dd = ctx->dev.take(plane_bytes(r));
            e = upload_plane(dr, r, s);
            if (e != hipSuccess) return hip_fail_locked(ctx, e, "reference upload");
            dev_frames[f].ref[c].data = dr;
            dev_frames[f].ref[c].stride = r.width;
            dev_frames[f].dst[c].data = dd;
            dev_frames[f].dst[c].stride = r.width;
        }
    if ((e = hipMemcpyAsync(d_frames, dev_frames.data(), fb, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(d_pus, pus, pb, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(d_items, items.data(), ib, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail_locked(ctx, e, "upload");
    e = rmh::launch_mc(d_frames, d_items, (uint32_t)items.size(), (uint32_t)n8, s);
    if (e != hipSuccess) return hip_fail_locked(ctx, e, "kernel launch");
    std::vector<void*> hdst;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            const rmh_plane& d = dev_frames[f].dst[c];
            void* h = ctx->host.take(plane_bytes(d));
            hdst.push_back(h);
            if ((e = hipMemcpyAsync(h, d.data, plane_bytes(d), hipMemcpyDeviceToHost, s)) != hipSuccess)
                return hip_fail_locked(ctx, e, "download");
        }
    rmh_status r = wait_stream(ctx, s);
    if (r != RMH_OK) return r;
    // Copy back only the samples some PU predicted; the rest keep the caller's values.
    for (size_t f = 0, k = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++, k++) {
            const rmh_plane& d = frames[f].dst[c];
            const uint8_t* src = (const uint8_t*)hdst[k];
            for (size_t i = 0; i < n_pu; i++) {
                if (pus[i].frame != f) continue;
                int sx = c ? 1 : 0;
                uint32_t x0 = pus[i].x >> sx, y0 = pus[i].y >> sx;
                uint32_t w = (1u << (pus[i].log2_w & 15)) >> sx, h = (1u << (pus[i].log2_h & 15)) >> sx;
                if (x0 >= d.width || y0 >= d.height) continue;
                if (x0 + w > d.width) w = d.width - x0;
                if (y0 + h > d.height) h = d.height - y0;
                for (uint32_t y = 0; y < h; y++)
                    memcpy((uint8_t*)d.data + ((size_t)(y0 + y) * d.stride + x0) * d.bytes_per_sample,
                           src + ((size_t)(y0 + y) * d.width + x0) * d.bytes_per_sample, (size_t)w * d.bytes_per_sample);
            }
        }
    return RMH_OK;
}

// ------------------------------------------------------------------ bi-prediction

RMH_EXPORT rmh_status rmh_mc_bi_plan_create(rmh_ctx* ctx, const rmh_mc_bi_frame* frames, size_t n_frames,
                                            const rmh_bi_pu_desc* pus, size_t n_pu, rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rmh_status v = validate_mc_bi(ctx, frames, n_frames, pus, n_pu);
        if (v != RMH_OK) return v;
    }
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    std::vector<rmh::McBiItem> items = mc_bi_items(frames, pus, n_pu);
    BiRefineWork rw = bi_refine_work(frames, pus, n_pu);
    if (items.size() > 0x7fffffffu || rw.items.size() > 0x7fffffffu)
        return fail(ctx, RMH_LIMIT_EXCEEDED, "too many inter work items");
    const size_t n8 = order_by_sample_size(items, frames);
    size_t fb = n_frames * sizeof(rmh_mc_bi_frame), ib = items.size() * sizeof(rmh::McBiItem);
    size_t jb = rw.jobs.size() * sizeof(rmh::DmvrJob), rb = rw.items.size() * sizeof(rmh::McRefItem);
    size_t mb = rw.mvs.size() * sizeof(int32_t);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kMc;
    p->n_items = items.size();
    p->n_items8 = n8;
    p->n_dmvr = rw.jobs.size();
    p->dmvr_lds = rw.lds;
    p->n_ref_items = rw.items.size();
    p->bytes = mc_bi_bytes(frames, pus, n_pu);
    const size_t o_items = aligned(fb), o_jobs = o_items + aligned(ib), o_ref = o_jobs + aligned(jb);
    const size_t o_mvs = o_ref + aligned(rb);
    hipError_t e = hipMalloc(&p->d_mem, o_mvs + aligned(mb) + 256);
    char* base = (char*)p->d_mem;
    if (e == hipSuccess && fb) e = hipMemcpy(base, frames, fb, hipMemcpyHostToDevice);
    if (e == hipSuccess && ib) e = hipMemcpy(base + o_items, items.data(), ib, hipMemcpyHostToDevice);
    if (e == hipSuccess && jb) e = hipMemcpy(base + o_jobs, rw.jobs.data(), jb, hipMemcpyHostToDevice);
    if (e == hipSuccess && rb) e = hipMemcpy(base + o_ref, rw.items.data(), rb, hipMemcpyHostToDevice);
    if (e == hipSuccess && mb) e = hipMemcpy(base + o_mvs, rw.mvs.data(), mb, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (p->d_mem) hipFree(p->d_mem);
        delete p;
        return fail(ctx, RMH_EXECUTION, std::string("bi-prediction plan upload failed: ") + hipGetErrorString(e));
    }
    p->d_bi_frames = (const rmh_mc_bi_frame*)base;
    p->d_bi_items = (const rmh::McBiItem*)(base + o_items);
    p->d_dmvr = (const rmh::DmvrJob*)(base + o_jobs);
    p->d_ref_items = (const rmh::McRefItem*)(base + o_ref);
    p->d_mvs = (int32_t*)(base + o_mvs);
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_mc_bi_batch(rmh_ctx* ctx, const rmh_mc_bi_frame* frames, size_t n_frames,
                                      const rmh_bi_pu_desc* pus, size_t n_pu) {
    if (!ctx) return RMH_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(ctx->mu);
    rmh_status v = validate_mc_bi(ctx, frames, n_frames, pus, n_pu);
    if (v != RMH_OK) return v;
    if (n_pu == 0) return RMH_OK;
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail_locked(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    std::vector<rmh::McBiItem> items = mc_bi_items(frames, pus, n_pu);
    BiRefineWork rw = bi_refine_work(frames, pus, n_pu);
    const size_t n8 = order_by_sample_size(items, frames);
    size_t fb = n_frames * sizeof(rmh_mc_bi_frame), ib = items.size() * sizeof(rmh::McBiItem);
    size_t jb = rw.jobs.size() * sizeof(rmh::DmvrJob), rb = rw.items.size() * sizeof(rmh::McRefItem);
    size_t mb = rw.mvs.size() * sizeof(int32_t);
    size_t need = aligned(fb) + aligned(ib) + aligned(jb) + aligned(rb) + aligned(mb), hneed = 0;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            need += 3 * aligned(plane_bytes(frames[f].ref0[c]));
            hneed += aligned(plane_bytes(frames[f].dst[c]));
        }
    hipError_t e = ctx->dev.reserve(need);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("device allocation: ") + hipGetErrorString(e));
    e = ctx->host.reserve(hneed);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("host staging: ") + hipGetErrorString(e));
    hipStream_t s = ctx->stream;
    rmh_mc_bi_frame* d_frames = (rmh_mc_bi_frame*)ctx->dev.take(fb);
    rmh::McBiItem* d_items = (rmh::McBiItem*)ctx->dev.take(ib);
    rmh::DmvrJob* d_jobs = (rmh::DmvrJob*)ctx->dev.take(jb);
    rmh::McRefItem* d_ref = (rmh::McRefItem*)ctx->dev.take(rb);
    int32_t* d_mvs = (int32_t*)ctx->dev.take(mb);
    std::vector<rmh_mc_bi_frame> dev_frames(frames, frames + n_frames);
    auto upload = [&](const rmh_plane& r, rmh_plane& d) -> hipError_t {
        void* dr = ctx->dev.take(plane_bytes(r));
        hipError_t u = upload_plane(dr, r, s);
        d.data = dr;
        d.stride = r.width;
        return u;
    };
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            if ((e = upload(frames[f].ref0[c], dev_frames[f].ref0[c])) != hipSuccess ||
                (e = upload(frames[f].ref1[c], dev_frames[f].ref1[c])) != hipSuccess)
                return hip_fail_locked(ctx, e, "reference upload");
            dev_frames[f].dst[c].data = ctx->dev.take(plane_bytes(frames[f].dst[c]));
            dev_frames[f].dst[c].stride = frames[f].dst[c].width;
        }
    if ((e = hipMemcpyAsync(d_frames, dev_frames.data(), fb, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (ib && (e = hipMemcpyAsync(d_items, items.data(), ib, hipMemcpyHostToDevice, s)) != hipSuccess) ||
        (jb && (e = hipMemcpyAsync(d_jobs, rw.jobs.data(), jb, hipMemcpyHostToDevice, s)) != hipSuccess) ||
        (rb && (e = hipMemcpyAsync(d_ref, rw.items.data(), rb, hipMemcpyHostToDevice, s)) != hipSuccess) ||
        (mb && (e = hipMemcpyAsync(d_mvs, rw.mvs.data(), mb, hipMemcpyHostToDevice, s)) != hipSuccess))
        return hip_fail_locked(ctx, e, "upload");
    e = rmh::launch_dmvr(d_frames, d_jobs, (uint32_t)rw.jobs.size(), rw.lds, d_mvs, s);
    if (e == hipSuccess) e = rmh::launch_mc_bi_refined(d_frames, d_ref, (uint32_t)rw.items.size(), d_mvs, s);
    if (e == hipSuccess) e = rmh::launch_mc_bi(d_frames, d_items, (uint32_t)items.size(), (uint32_t)n8, s);
    if (e != hipSuccess) return hip_fail_locked(ctx, e, "kernel launch");
    std::vector<void*> hdst;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            const rmh_plane& d = dev_frames[f].dst[c];
            void* h = ctx->host.take(plane_bytes(d));
            hdst.push_back(h);
            if ((e = hipMemcpyAsync(h, d.data, plane_bytes(d), hipMemcpyDeviceToHost, s)) != hipSuccess)
                return hip_fail_locked(ctx, e, "download");
        }
    rmh_status r = wait_stream(ctx, s);
    if (r != RMH_OK) return r;
    // Copy back only the samples some PU predicted; the rest keep the caller's values.
    for (size_t f = 0, k = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++, k++) {
            const rmh_plane& d = frames[f].dst[c];
            const uint8_t* src = (const uint8_t*)hdst[k];
            for (size_t i = 0; i < n_pu; i++) {
                if (pus[i].frame != f) continue;
                int sx = c ? 1 : 0;
                uint32_t x0 = pus[i].x >> sx, y0 = pus[i].y >> sx;
                uint32_t w = (1u << (pus[i].log2_w & 15)) >> sx, h = (1u << (pus[i].log2_h & 15)) >> sx;
                if (x0 >= d.width || y0 >= d.height) continue;
                if (x0 + w > d.width) w = d.width - x0;
                if (y0 + h > d.height) h = d.height - y0;
                for (uint32_t y = 0; y < h; y++)
                    memcpy((uint8_t*)d.data + ((size_t)(y0 + y) * d.stride + x0) * d.bytes_per_sample,
                           src + ((size_t)(y0 + y) * d.width + x0) * d.bytes_per_sample, (size_t)w * d.bytes_per_sample);
            }
        }
    return RMH_OK;
}

// ------------------------------------------------------------------ deblock

RMH_EXPORT rmh_status rmh_deblock_plan_create(rmh_ctx* ctx, const rmh_dbk_frame* frames, size_t n_frames, rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rmh_status v = validate_dbk(ctx, frames, n_frames);
        if (v != RMH_OK) return v;
    }
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kDeblock;
    p->n_dbk = n_frames;
    const TileTable tt = tile_table(frames, n_frames);
    p->total_tiles = tt.total;
    p->tiles8 = tt.tiles8;
    for (size_t f = 0; f < n_frames; f++) {
        uint32_t cells = frames[f].grid_w * frames[f].grid_h;
        p->max_rows = std::max(p->max_rows, frames[f].planes[0].height);
        for (int c = 0; c < 3; c++) p->bytes += 2 * plane_bytes(frames[f].planes[c]);
        p->bytes += 4ull * cells;
    }
    size_t fb = n_frames * sizeof(rmh_dbk_frame), sb = n_frames * 3 * sizeof(rmh_plane);
    size_t tb = tt.words.size() * sizeof(uint32_t), shb = shadow_bytes(frames, n_frames);
    hipError_t e = hipMalloc(&p->d_mem, aligned(fb) + aligned(sb) + aligned(tb) + shb + 256);
    char* base = (char*)p->d_mem;
    if (e == hipSuccess) {
        std::vector<rmh_plane> sh = shadow_layout(frames, n_frames, base + aligned(fb) + aligned(sb) + aligned(tb));
        if (fb) e = hipMemcpy(base, frames, fb, hipMemcpyHostToDevice);
        if (e == hipSuccess && sb) e = hipMemcpy(base + aligned(fb), sh.data(), sb, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(base + aligned(fb) + aligned(sb), tt.words.data(), tb, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        if (p->d_mem) hipFree(p->d_mem);
        delete p;
        return fail(ctx, RMH_EXECUTION, std::string("deblock plan upload failed: ") + hipGetErrorString(e));
    }
    p->d_dbk = (const rmh_dbk_frame*)base;
    p->d_shadows = (const rmh_plane*)(base + aligned(fb));
    p->d_tile_start = (const uint32_t*)(base + aligned(fb) + aligned(sb));
    p->d_frame_of = p->d_tile_start + n_frames + 1;
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_deblock_to_plan_create(rmh_ctx* ctx, const rmh_dbk_frame* frames, size_t n_frames,
                                                 const rmh_plane* dst, rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rmh_status v = validate_dbk(ctx, frames, n_frames, false);
        if (v != RMH_OK) return v;
        if (n_frames && !dst) return fail_locked(ctx, RMH_INVALID_INPUT, "null destination planes");
        std::vector<PlaneSpan> spans;
        for (size_t f = 0; f < n_frames; f++)
            for (int c = 0; c < 3; c++) {
                add_span(spans, frames[f].planes[c], false);
                if (plane_ok(dst[3 * f + c])) add_span(spans, dst[3 * f + c], true);
            }
        if (written_overlap(spans))
            return fail_locked(ctx, RMH_INVALID_INPUT, "a destination plane overlaps another plane of the batch");
        for (size_t f = 0; f < n_frames; f++)
            for (int c = 0; c < 3; c++) {
                const rmh_plane &a = frames[f].planes[c], &b = dst[3 * f + c];
                if (!plane_ok(b) || b.width != a.width || b.height != a.height ||
                    b.bytes_per_sample != a.bytes_per_sample || b.bit_depth != a.bit_depth)
                    return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": invalid destination plane");
                for (int k = 0; k < 3; k++) {
                    const rmh_plane& q = frames[f].planes[k];
                    const uintptr_t qa = (uintptr_t)q.data, ba = (uintptr_t)b.data;
                    const uint64_t qn = (uint64_t)q.stride * q.height * q.bytes_per_sample;
                    const uint64_t bn = (uint64_t)b.stride * b.height * b.bytes_per_sample;
                    if (qa < ba + bn && ba < qa + qn)
                        return fail_locked(ctx, RMH_INVALID_INPUT, "frame " + std::to_string(f) + ": destination overlaps a source plane");
                }
            }
    }
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kDeblock;
    p->n_dbk = n_frames;
    p->dbk_copy = false;
    const TileTable tt = tile_table(frames, n_frames);
    p->total_tiles = tt.total;
    p->tiles8 = tt.tiles8;
    std::vector<rmh_dbk_frame> outf(frames, frames + n_frames);
    std::vector<rmh_plane> srcs(n_frames * 3);
    for (size_t f = 0; f < n_frames; f++) {
        uint32_t cells = frames[f].grid_w * frames[f].grid_h;
        p->max_rows = std::max(p->max_rows, frames[f].planes[0].height);
        for (int c = 0; c < 3; c++) {
            p->bytes += 2 * plane_bytes(frames[f].planes[c]);
            srcs[3 * f + c] = frames[f].planes[c];
            outf[f].planes[c] = dst[3 * f + c];
        }
        p->bytes += 4ull * cells;
    }
    size_t fb = n_frames * sizeof(rmh_dbk_frame), sb = n_frames * 3 * sizeof(rmh_plane);
    size_t tb = tt.words.size() * sizeof(uint32_t);
    hipError_t e = hipMalloc(&p->d_mem, aligned(fb) + aligned(sb) + aligned(tb) + 256);
    char* base = (char*)p->d_mem;
    if (e == hipSuccess) {
        if (fb) e = hipMemcpy(base, outf.data(), fb, hipMemcpyHostToDevice);
        if (e == hipSuccess && sb) e = hipMemcpy(base + aligned(fb), srcs.data(), sb, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(base + aligned(fb) + aligned(sb), tt.words.data(), tb, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        if (p->d_mem) hipFree(p->d_mem);
        delete p;
        return fail(ctx, RMH_EXECUTION, std::string("deblock plan upload failed: ") + hipGetErrorString(e));
    }
    p->d_dbk = (const rmh_dbk_frame*)base;
    p->d_shadows = (const rmh_plane*)(base + aligned(fb));
    p->d_tile_start = (const uint32_t*)(base + aligned(fb) + aligned(sb));
    p->d_frame_of = p->d_tile_start + n_frames + 1;
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_deblock_plan_run(rmh_plan* plan, void* stream) {
    if (!plan || plan->kind != kDeblock) return RMH_INVALID_INPUT;
    if (plan->n_dbk == 0) return RMH_OK;
    return run_plan(plan, nullptr, nullptr, 0, 0, stream);
}

// ------------------------------------------------------------------ edge map

namespace {

// Edge-map frames: grids that match the visible size, a CTB size the
// decoder allows, and (device_edges) 4-byte aligned edge buffers.
rmh_status validate_edge(rmh_ctx* ctx, const rmh_edge_frame* frames, size_t n_frames, bool device_edges,
                         uint64_t* total) {
    if (n_frames && !frames) return fail(ctx, RMH_INVALID_INPUT, "null frame array");
    if (n_frames > 65535) return fail(ctx, RMH_LIMIT_EXCEEDED, "more than 65535 frames");
    *total = 0;
    for (size_t f = 0; f < n_frames; f++) {
        const rmh_edge_frame& e = frames[f];
        const std::string tag = "frame " + std::to_string(f) + ": ";
        if (!e.cells || !e.edges) return fail(ctx, RMH_INVALID_INPUT, tag + "null cells or edges");
        if (e.grid_w == 0 || e.grid_h == 0 || e.grid_w > 16384 || e.grid_h > 16384 || e.reserved)
            return fail(ctx, RMH_INVALID_INPUT, tag + "grid size");
        if (e.vis_w > 4 * e.grid_w || e.vis_w + 3 < 4 * e.grid_w || e.vis_h > 4 * e.grid_h || e.vis_h + 3 < 4 * e.grid_h)
            return fail(ctx, RMH_INVALID_INPUT, tag + "visible size does not match the grid");
        if (e.ctb_size < 16 || e.ctb_size > 128 || (e.ctb_size & (e.ctb_size - 1)))
            return fail(ctx, RMH_INVALID_INPUT, tag + "CTB size");
        if (device_edges && ((uintptr_t)e.edges & 3)) return fail(ctx, RMH_INVALID_INPUT, tag + "edges not 4-byte aligned");
        *total += (uint64_t)e.grid_w * e.grid_h;
    }
    if (*total > 0xffffffffull) return fail(ctx, RMH_LIMIT_EXCEEDED, "more than 2^32 cells");
    return RMH_OK;
}

std::vector<uint32_t> cell_starts(const rmh_edge_frame* frames, size_t n_frames) {
    std::vector<uint32_t> v(n_frames + 1, 0);
    for (size_t f = 0; f < n_frames; f++) v[f + 1] = v[f] + frames[f].grid_w * frames[f].grid_h;
    return v;
}

}  // namespace

RMH_EXPORT rmh_status rmh_edge_map_plan_create(rmh_ctx* ctx, const rmh_edge_frame* frames, size_t n_frames,
                                               rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    uint64_t total = 0;
    rmh_status v = validate_edge(ctx, frames, n_frames, true, &total);
    if (v != RMH_OK) return v;
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kEdge;
    p->n_edge = n_frames;
    p->total_cells = (uint32_t)total;
    p->bytes = total * (sizeof(rmh_edge_cell) + 4);
    const std::vector<uint32_t> starts = cell_starts(frames, n_frames);
    const size_t fb = aligned(n_frames * sizeof(rmh_edge_frame)), sb = aligned(starts.size() * 4);
    hipError_t e = hipMalloc(&p->d_mem, fb + sb);
    char* base = (char*)p->d_mem;
    if (e == hipSuccess && n_frames) e = hipMemcpy(base, frames, n_frames * sizeof(rmh_edge_frame), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(base + fb, starts.data(), starts.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (p->d_mem) hipFree(p->d_mem);
        delete p;
        return fail(ctx, RMH_EXECUTION, std::string("edge-map plan upload failed: ") + hipGetErrorString(e));
    }
    p->d_edge_frames = (const rmh_edge_frame*)base;
    p->d_cell_start = (const uint32_t*)(base + fb);
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_edge_map_plan_run(rmh_plan* plan, void* stream) {
    if (!plan || plan->kind != kEdge) return RMH_INVALID_INPUT;
    if (plan->total_cells == 0) return RMH_OK;
    return run_plan(plan, nullptr, nullptr, 0, 0, stream);
}

RMH_EXPORT rmh_status rmh_edge_map(rmh_ctx* ctx, const rmh_edge_frame* frames, size_t n_frames) {
    if (!ctx) return RMH_INVALID_INPUT;
    uint64_t total = 0;
    rmh_status v = validate_edge(ctx, frames, n_frames, false, &total);  // reports through fail(), before the lock
    if (v != RMH_OK) return v;
    if (total == 0) return RMH_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail_locked(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    const std::vector<uint32_t> starts = cell_starts(frames, n_frames);
    const size_t fb = aligned(n_frames * sizeof(rmh_edge_frame)), sb = aligned(starts.size() * 4);
    const size_t cb = aligned(total * sizeof(rmh_edge_cell)), eb = aligned(total * 4);
    hipError_t e = ctx->dev.reserve(fb + sb + cb + eb);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("device allocation: ") + hipGetErrorString(e));
    e = ctx->host.reserve(eb);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("host staging: ") + hipGetErrorString(e));
    rmh_edge_frame* d_frames = (rmh_edge_frame*)ctx->dev.take(fb);
    uint32_t* d_starts = (uint32_t*)ctx->dev.take(sb);
    rmh_edge_cell* d_cells = (rmh_edge_cell*)ctx->dev.take(cb);
    uint8_t* d_edges = (uint8_t*)ctx->dev.take(eb);
    uint8_t* h_edges = (uint8_t*)ctx->host.take(eb);
    std::vector<rmh_edge_frame> dev(frames, frames + n_frames);
    hipStream_t s = ctx->stream;
    for (size_t f = 0; f < n_frames && e == hipSuccess; f++) {
        const size_t n = (size_t)frames[f].grid_w * frames[f].grid_h;
        dev[f].cells = d_cells + starts[f];
        dev[f].edges = d_edges + 4 * (size_t)starts[f];
        e = hipMemcpyAsync((void*)dev[f].cells, frames[f].cells, n * sizeof(rmh_edge_cell), hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(d_frames, dev.data(), n_frames * sizeof(rmh_edge_frame), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_starts, starts.data(), starts.size() * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail_locked(ctx, e, "upload");
    if ((e = rmh::launch_edge_map(d_frames, d_starts, (uint32_t)n_frames, (uint32_t)total, s)) != hipSuccess)
        return hip_fail_locked(ctx, e, "kernel launch");
    if ((e = hipMemcpyAsync(h_edges, d_edges, total * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return hip_fail_locked(ctx, e, "download");
    rmh_status r = wait_stream(ctx, s);
    if (r != RMH_OK) return r;
    for (size_t f = 0; f < n_frames; f++)
        memcpy(frames[f].edges, h_edges + 4 * (size_t)starts[f], 4 * (size_t)frames[f].grid_w * frames[f].grid_h);
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_deblock_noalloc(rmh_ctx* ctx, const rmh_dbk_frame* frames, size_t n_frames, void* stream) {
    if (!ctx) return RMH_INVALID_INPUT;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rmh_status v = validate_dbk(ctx, frames, n_frames);
        if (v != RMH_OK) return v;
    }
    if (n_frames == 0) return RMH_OK;
    // the tables: frames, shadow plane records, tile running sum + frame order
    const size_t fb = aligned(n_frames * sizeof(rmh_dbk_frame)), sb = aligned(n_frames * 3 * sizeof(rmh_plane));
    const size_t tb = aligned((2 * n_frames + 1) * sizeof(uint32_t));
    if (fb + sb + tb > kScratchTableBytes)
        return fail(ctx, RMH_LIMIT_EXCEEDED, "too many frames for the no-allocation scratch tables");
    if (shadow_bytes(frames, n_frames) > kScratchBytes - kScratchTableBytes)
        return fail(ctx, RMH_LIMIT_EXCEEDED,
                    "planes exceed the " + std::to_string((kScratchBytes - kScratchTableBytes) >> 10) +
                        " KiB of no-allocation scratch for plane copies");
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    char* base = (char*)ctx->scratch;
    const TileTable tt = tile_table(frames, n_frames);
    const std::vector<rmh_plane> shadows = shadow_layout(frames, n_frames, base + kScratchTableBytes);
    uint32_t max_rows = 0;
    for (size_t f = 0; f < n_frames; f++) max_rows = std::max(max_rows, frames[f].planes[0].height);
    static_assert(sizeof(rmh_dbk_frame) % 4 == 0 && sizeof(rmh_plane) % 4 == 0, "tables are written as words");
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = write_words(base, frames, n_frames * sizeof(rmh_dbk_frame), s);
    if (e == hipSuccess) e = write_words(base + fb, shadows.data(), shadows.size() * sizeof(rmh_plane), s);
    if (e == hipSuccess) e = write_words(base + fb + sb, tt.words.data(), tt.words.size() * sizeof(uint32_t), s);
    const uint32_t* d_words = (const uint32_t*)(base + fb + sb);
    if (e == hipSuccess)
        e = rmh::launch_deblock((const rmh_dbk_frame*)base, (const rmh_plane*)(base + fb), d_words, d_words + n_frames + 1,
                                (uint32_t)n_frames, tt.total, tt.tiles8, max_rows, true, s);
    if (e != hipSuccess) return fail(ctx, RMH_EXECUTION, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_deblock(rmh_ctx* ctx, const rmh_dbk_frame* frames, size_t n_frames) {
    if (!ctx) return RMH_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(ctx->mu);
    rmh_status v = validate_dbk(ctx, frames, n_frames);
    if (v != RMH_OK) return v;
    if (n_frames == 0) return RMH_OK;
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail_locked(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    size_t fb = n_frames * sizeof(rmh_dbk_frame), sb = n_frames * 3 * sizeof(rmh_plane);
    const TileTable tt = tile_table(frames, n_frames);
    const size_t tb = tt.words.size() * sizeof(uint32_t);
    size_t need = aligned(fb) + aligned(sb) + aligned(tb) + shadow_bytes(frames, n_frames), hneed = 0;
    uint32_t max_rows = 0;
    for (size_t f = 0; f < n_frames; f++) {
        uint32_t cells = frames[f].grid_w * frames[f].grid_h;
        max_rows = std::max(max_rows, frames[f].planes[0].height);
        need += aligned(4ull * cells);
        for (int c = 0; c < 3; c++) {
            need += aligned(plane_bytes(frames[f].planes[c]));
            hneed += aligned(plane_bytes(frames[f].planes[c]));
        }
    }
    hipError_t e = ctx->dev.reserve(need);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("device allocation: ") + hipGetErrorString(e));
    e = ctx->host.reserve(hneed);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("host staging: ") + hipGetErrorString(e));
    hipStream_t s = ctx->stream;
    rmh_dbk_frame* d_frames = (rmh_dbk_frame*)ctx->dev.take(fb);
    rmh_plane* d_shadows = (rmh_plane*)ctx->dev.take(sb);
    uint32_t* d_starts = (uint32_t*)ctx->dev.take(tb);
    if ((e = hipMemcpyAsync(d_starts, tt.words.data(), tb, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail_locked(ctx, e, "tile table upload");
    const std::vector<rmh_plane> sh = shadow_layout(frames, n_frames, (char*)ctx->dev.take(shadow_bytes(frames, n_frames)));
    if ((e = hipMemcpyAsync(d_shadows, sh.data(), sb, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail_locked(ctx, e, "shadow upload");
    std::vector<rmh_dbk_frame> df(frames, frames + n_frames);
    for (size_t f = 0; f < n_frames; f++) {
        size_t eb = 4ull * frames[f].grid_w * frames[f].grid_h;
        void* de = ctx->dev.take(eb);
        if ((e = hipMemcpyAsync(de, frames[f].edges, eb, hipMemcpyHostToDevice, s)) != hipSuccess)
            return hip_fail_locked(ctx, e, "edge upload");
        df[f].edges = (const uint8_t*)de;
        for (int c = 0; c < 3; c++) {
            const rmh_plane& p = frames[f].planes[c];
            void* dp = ctx->dev.take(plane_bytes(p));
            e = upload_plane(dp, p, s);
            if (e != hipSuccess) return hip_fail_locked(ctx, e, "plane upload");
            df[f].planes[c].data = dp;
            df[f].planes[c].stride = p.width;
        }
    }
    if ((e = hipMemcpyAsync(d_frames, df.data(), fb, hipMemcpyHostToDevice, s)) != hipSuccess) return hip_fail_locked(ctx, e, "upload");
    e = rmh::launch_deblock(d_frames, d_shadows, d_starts, d_starts + n_frames + 1, (uint32_t)n_frames, tt.total,
                            tt.tiles8, max_rows, true, s);
    if (e != hipSuccess) return hip_fail_locked(ctx, e, "kernel launch");
    std::vector<void*> hp;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            const rmh_plane& p = df[f].planes[c];
            void* h = ctx->host.take(plane_bytes(p));
            hp.push_back(h);
            if ((e = hipMemcpyAsync(h, p.data, plane_bytes(p), hipMemcpyDeviceToHost, s)) != hipSuccess)
                return hip_fail_locked(ctx, e, "download");
        }
    rmh_status r = wait_stream(ctx, s);
    if (r != RMH_OK) return r;
    for (size_t f = 0, k = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++, k++) {
            const rmh_plane& p = frames[f].planes[c];
            for (uint32_t y = 0; y < p.height; y++)
                memcpy((uint8_t*)p.data + (size_t)y * p.stride * p.bytes_per_sample,
                       (const uint8_t*)hp[k] + (size_t)y * p.width * p.bytes_per_sample, (size_t)p.width * p.bytes_per_sample);
        }
    return RMH_OK;
}

// ---------------------------------------------------------------------- sao

RMH_EXPORT rmh_status rmh_sao_plan_create(rmh_ctx* ctx, const rmh_sao_frame* frames, size_t n_frames, rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        rmh_status v = validate_sao(ctx, frames, n_frames, false);
        if (v != RMH_OK) return v;
    }
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    std::vector<const rmh_sao_param*> prm(n_frames);
    std::vector<std::array<rmh_plane, 6>> planes(n_frames);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kSao;
    for (size_t f = 0; f < n_frames; f++) {
        prm[f] = frames[f].params;
        for (int c = 0; c < 3; c++) {
            planes[f][c] = frames[f].src[c];
            planes[f][3 + c] = frames[f].dst[c];
            p->bytes += 2 * plane_bytes(frames[f].src[c]);
        }
        p->bytes += (uint64_t)frames[f].ctb_cols * frames[f].ctb_rows * 3 * sizeof(rmh_sao_param);
    }
    const SaoJobs sj = sao_jobs(frames, n_frames, prm, planes);
    p->n_sao_jobs = (uint32_t)sj.jobs.size();
    memcpy(p->sao_range, sj.range, sizeof sj.range);
    if (!sj.jobs.empty()) {
        const size_t jb = sj.jobs.size() * sizeof(rmh::SaoJob), bb = sj.block_job.size() * sizeof(uint32_t);
        hipError_t e = hipMalloc(&p->d_mem, aligned(jb) + bb);
        if (e == hipSuccess) e = hipMemcpy(p->d_mem, sj.jobs.data(), jb, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy((char*)p->d_mem + aligned(jb), sj.block_job.data(), bb, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (p->d_mem) hipFree(p->d_mem);
            delete p;
            return fail(ctx, RMH_EXECUTION, std::string("sao plan upload failed: ") + hipGetErrorString(e));
        }
        p->d_sao_jobs = (const rmh::SaoJob*)p->d_mem;
        p->d_sao_block_job = (const uint32_t*)((char*)p->d_mem + aligned(jb));
    }
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_sao_plan_run(rmh_plan* plan, void* stream) {
    if (!plan || plan->kind != kSao) return RMH_INVALID_INPUT;
    if (plan->n_sao_jobs == 0) return RMH_OK;
    return run_plan(plan, nullptr, nullptr, 0, 0, stream);
}

RMH_EXPORT rmh_status rmh_sao(rmh_ctx* ctx, const rmh_sao_frame* frames, size_t n_frames) {
    if (!ctx) return RMH_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(ctx->mu);
    rmh_status v = validate_sao(ctx, frames, n_frames, true);
    if (v != RMH_OK) return v;
    if (n_frames == 0) return RMH_OK;
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail_locked(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    size_t need = 0, hneed = 0;
    for (size_t f = 0; f < n_frames; f++) {
        need += aligned((size_t)frames[f].ctb_cols * frames[f].ctb_rows * 3 * sizeof(rmh_sao_param));
        for (int c = 0; c < 3; c++) {
            need += 2 * aligned(plane_bytes(frames[f].src[c]));
            hneed += aligned(plane_bytes(frames[f].src[c]));
        }
    }
    need += aligned(n_frames * 3 * sizeof(rmh::SaoJob));
    {
        // the work table's size (addresses do not change it)
        std::vector<const rmh_sao_param*> p0(n_frames, nullptr);
        std::vector<std::array<rmh_plane, 6>> q0(n_frames);
        for (size_t f = 0; f < n_frames; f++)
            for (int c = 0; c < 3; c++) q0[f][c] = q0[f][3 + c] = frames[f].src[c];
        need += aligned(sao_jobs(frames, n_frames, p0, q0).block_job.size() * sizeof(uint32_t));
    }
    hipError_t e = ctx->dev.reserve(need);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("device allocation: ") + hipGetErrorString(e));
    e = ctx->host.reserve(hneed);
    if (e != hipSuccess) return fail_locked(ctx, RMH_LIMIT_EXCEEDED, std::string("host staging: ") + hipGetErrorString(e));
    hipStream_t s = ctx->stream;
    std::vector<const rmh_sao_param*> prm(n_frames);
    std::vector<std::array<rmh_plane, 6>> planes(n_frames);
    for (size_t f = 0; f < n_frames; f++) {
        const size_t pb = (size_t)frames[f].ctb_cols * frames[f].ctb_rows * 3 * sizeof(rmh_sao_param);
        void* dp = ctx->dev.take(pb);
        if ((e = hipMemcpyAsync(dp, frames[f].params, pb, hipMemcpyHostToDevice, s)) != hipSuccess)
            return hip_fail_locked(ctx, e, "parameter upload");
        prm[f] = (const rmh_sao_param*)dp;
        for (int c = 0; c < 3; c++) {
            const rmh_plane& src = frames[f].src[c];
            rmh_plane d = src, o = src;
            d.data = ctx->dev.take(plane_bytes(src));
            d.stride = src.width;
            o.data = ctx->dev.take(plane_bytes(src));
            o.stride = src.width;
            e = upload_plane(d.data, src, s);
            if (e != hipSuccess) return hip_fail_locked(ctx, e, "plane upload");
            planes[f][c] = d;
            planes[f][3 + c] = o;
        }
    }
    const SaoJobs sj = sao_jobs(frames, n_frames, prm, planes);
    rmh::SaoJob* d_jobs = (rmh::SaoJob*)ctx->dev.take(sj.jobs.size() * sizeof(rmh::SaoJob));
    uint32_t* d_bj = (uint32_t*)ctx->dev.take(sj.block_job.size() * sizeof(uint32_t));
    if ((e = hipMemcpyAsync(d_jobs, sj.jobs.data(), sj.jobs.size() * sizeof(rmh::SaoJob), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(d_bj, sj.block_job.data(), sj.block_job.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                            s)) != hipSuccess)
        return hip_fail_locked(ctx, e, "upload");
    e = rmh::launch_sao(d_jobs, d_bj, sj.range, s);
    if (e != hipSuccess) return hip_fail_locked(ctx, e, "kernel launch");
    std::vector<void*> hp;
    for (size_t f = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++) {
            const rmh_plane& o = planes[f][3 + c];
            void* h = ctx->host.take(plane_bytes(o));
            hp.push_back(h);
            if ((e = hipMemcpyAsync(h, o.data, plane_bytes(o), hipMemcpyDeviceToHost, s)) != hipSuccess)
                return hip_fail_locked(ctx, e, "download");
        }
    rmh_status r = wait_stream(ctx, s);
    if (r != RMH_OK) return r;
    for (size_t f = 0, k = 0; f < n_frames; f++)
        for (int c = 0; c < 3; c++, k++) {
            const rmh_plane& d = frames[f].dst[c];
            for (uint32_t y = 0; y < d.height; y++)
                memcpy((uint8_t*)d.data + (size_t)y * d.stride * d.bytes_per_sample,
                       (const uint8_t*)hp[k] + (size_t)y * d.width * d.bytes_per_sample, (size_t)d.width * d.bytes_per_sample);
        }
    return RMH_OK;
}

// ----------------------------------------------------------- forward (encoder)

RMH_EXPORT rmh_status rmh_forward_plan_create(rmh_ctx* ctx, const rmh_tu_desc* tus, size_t n_tu, size_t n_coeff,
                                              rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    if (n_tu && !tus) return fail(ctx, RMH_INVALID_INPUT, "null TU array");
    // the inverse plan checks every TU as a decoder TU; the forward flag and
    // joint Cb-Cr (a decoder-side derivation) are checked here
    std::vector<rmh_tu_desc> inv(tus, tus + n_tu);
    for (size_t i = 0; i < n_tu; i++) {
        if (inv[i].joint) return fail(ctx, RMH_INVALID_INPUT, "TU " + std::to_string(i) + ": joint Cb-Cr in a forward plan");
        inv[i].flags &= (uint8_t)~RMH_TU_QUANT_INTER;
    }
    rmh_plan* sub = nullptr;
    rmh_status st = create_residual_plan(ctx, inv.data(), n_tu, n_coeff, nullptr, nullptr, 0, &sub);
    if (st != RMH_OK) return st;
    DeviceGuard dg(ctx->device);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kForward;
    p->n_tu = n_tu;
    p->n_coeff = n_coeff;
    p->inverse = sub;
    for (size_t i = 0; i < n_tu; i++) p->bytes += 4ull << (tus[i].log2_w + tus[i].log2_h);  // residual in, levels out
    if (n_tu) {
        hipError_t e = hipMalloc(&p->d_mem, n_tu * sizeof(rmh_tu_desc));
        if (e == hipSuccess) e = hipMemcpy(p->d_mem, tus, n_tu * sizeof(rmh_tu_desc), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            rmh_plan_destroy(p);
            return fail(ctx, RMH_EXECUTION, std::string("forward plan upload failed: ") + hipGetErrorString(e));
        }
        p->d_tus = (const rmh_tu_desc*)p->d_mem;
    }
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_forward_plan_run(rmh_plan* plan, const int16_t* d_residual, int16_t* d_levels, void* stream) {
    if (!plan || plan->kind != kForward) return RMH_INVALID_INPUT;
    if (plan->n_tu == 0) return RMH_OK;
    if (!d_residual || !d_levels) return fail(plan->ctx, RMH_INVALID_INPUT, "null device buffer");
    return run_plan(plan, d_residual, d_levels, 0, plan->n_tu, stream);
}

RMH_EXPORT rmh_status rmh_forward_plan_evaluate(rmh_plan* plan, const int16_t* d_residual, const int16_t* d_pred,
                                                int16_t* d_levels, int16_t* d_recon, uint64_t* d_sse, void* stream) {
    if (!plan || plan->kind != kForward) return RMH_INVALID_INPUT;
    if (plan->n_tu == 0) return RMH_OK;
    if (!d_residual || !d_levels || !d_recon || !d_sse) return fail(plan->ctx, RMH_INVALID_INPUT, "null device buffer");
    rmh_status st = run_plan(plan, d_residual, d_levels, 0, plan->n_tu, stream);
    if (st == RMH_OK) st = run_plan(plan->inverse, d_levels, d_recon, 0, plan->n_tu, stream);
    if (st != RMH_OK) return st;
    DeviceGuard dg(plan->ctx->device);
    rmh::SseArgs a{};
    a.tus = plan->d_tus;
    a.residual = d_residual;
    a.pred = d_pred;
    a.recon = d_recon;
    a.sse = d_sse;
    a.n_tu = (uint32_t)plan->n_tu;
    hipError_t e = rmh::launch_sse(a, (hipStream_t)stream);
    if (e != hipSuccess) return fail(plan->ctx, RMH_EXECUTION, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return RMH_OK;
}

// ------------------------------------------------------------- lmcs, alf

namespace {

// Device blob of a loop-filter plan: parts appended at 256-byte offsets.
struct Blob {
    std::vector<unsigned char> bytes;
    size_t add(const void* p, size_t n) {
        const size_t off = aligned(bytes.size());
        bytes.resize(off + n);
        if (n) memcpy(bytes.data() + off, p, n);
        return off;
    }
};

// Uploads the blob into the plan's device memory; false on failure.
bool upload_blob(rmh_plan* p, const Blob& b, std::string* err) {
    if (b.bytes.empty()) return true;
    hipError_t e = hipMalloc(&p->d_mem, b.bytes.size());
    if (e == hipSuccess) e = hipMemcpy(p->d_mem, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        *err = hipGetErrorString(e);
        if (p->d_mem) hipFree(p->d_mem);
        p->d_mem = nullptr;
        return false;
    }
    return true;
}

bool same_shape(const rmh_plane& a, const rmh_plane& b) {
    return a.width == b.width && a.height == b.height && a.bytes_per_sample == b.bytes_per_sample &&
           a.bit_depth == b.bit_depth;
}

rmh_status validate_lmcs(rmh_ctx* ctx, const rmh_lmcs_frame* frames, size_t n_frames) {
    if (n_frames && !frames) return fail(ctx, RMH_INVALID_INPUT, "null frame array");
    if (n_frames > 65535) return fail(ctx, RMH_LIMIT_EXCEEDED, "more than 65535 frames");
    std::vector<PlaneSpan> spans;
    for (size_t f = 0; f < n_frames; f++) {
        const rmh_lmcs_frame& d = frames[f];
        const std::string tag = "frame " + std::to_string(f) + ": ";
        if (!plane_ok(d.src) || !plane_ok(d.dst) || !same_shape(d.src, d.dst))
            return fail(ctx, RMH_INVALID_INPUT, tag + "invalid or mismatched src/dst plane");
        if (d.direction > RMH_LMCS_FORWARD || d.aps.reserved) return fail(ctx, RMH_INVALID_INPUT, tag + "direction or APS");
        if (d.src.data == d.dst.data) {
            if (d.src.stride != d.dst.stride) return fail(ctx, RMH_INVALID_INPUT, tag + "in place with two strides");
        } else {
            add_span(spans, d.src, false);
        }
        add_span(spans, d.dst, true);
    }
    if (written_overlap(spans)) return fail(ctx, RMH_INVALID_INPUT, "a dst plane overlaps another plane of the batch");
    return RMH_OK;
}

bool alf_aps_ok(const rmh_alf_aps* a, bool luma) {
    if (!a) return true;
    if (a->n_chroma > 8 || a->n_cc_cb > 4 || a->n_cc_cr > 4 || a->n_luma > 25 || (luma && a->n_luma == 0)) return false;
    for (int f = 0; f < 25; f++)
        for (int k = 0; k < 12; k++)
            if (a->luma_coeff[f][k] < -128 || a->luma_coeff[f][k] > 128) return false;
    for (int f = 0; f < 8; f++)
        for (int k = 0; k < 6; k++)
            if (a->chroma_coeff[f][k] < -128 || a->chroma_coeff[f][k] > 128) return false;
    for (int f = 0; f < 4; f++)
        for (int k = 0; k < 7; k++)
            if (a->cc_cb[f][k] < -64 || a->cc_cb[f][k] > 64 || a->cc_cr[f][k] < -64 || a->cc_cr[f][k] > 64) return false;
    return true;
}

rmh_status validate_alf(rmh_ctx* ctx, const rmh_alf_frame* frames, size_t n_frames) {
    if (n_frames && !frames) return fail(ctx, RMH_INVALID_INPUT, "null frame array");
    if (n_frames > 65535) return fail(ctx, RMH_LIMIT_EXCEEDED, "more than 65535 frames");
    std::vector<PlaneSpan> spans;
    for (size_t f = 0; f < n_frames; f++) {
        const rmh_alf_frame& d = frames[f];
        const std::string tag = "frame " + std::to_string(f) + ": ";
        for (int c = 0; c < 3; c++) {
            if (!plane_ok(d.src[c]) || !plane_ok(d.dst[c]) || !same_shape(d.src[c], d.dst[c]) ||
                d.src[c].bit_depth != d.src[0].bit_depth || d.src[c].bytes_per_sample != d.src[0].bytes_per_sample)
                return fail(ctx, RMH_INVALID_INPUT, tag + "invalid or mismatched src/dst plane");
            add_span(spans, d.src[c], false);
            add_span(spans, d.dst[c], true);
        }
        if (d.src[1].width != (d.src[0].width + 1) / 2 || d.src[1].height != (d.src[0].height + 1) / 2 ||
            d.src[2].width != d.src[1].width || d.src[2].height != d.src[1].height)
            return fail(ctx, RMH_INVALID_INPUT, tag + "chroma planes are not 4:2:0");
        if (d.ctb_size != 16 && d.ctb_size != 32 && d.ctb_size != 64 && d.ctb_size != 128)
            return fail(ctx, RMH_INVALID_INPUT, tag + "CTB size");
        if (d.vis_w == 0 || d.vis_h == 0 || d.vis_w > d.src[0].width || d.vis_h > d.src[0].height || d.reserved)
            return fail(ctx, RMH_INVALID_INPUT, tag + "visible size");
        if (!d.ctbs || d.ctb_cols < (d.vis_w + d.ctb_size - 1) / d.ctb_size ||
            d.ctb_rows < (d.vis_h + d.ctb_size - 1) / d.ctb_size)
            return fail(ctx, RMH_INVALID_INPUT, tag + "CTB switch grid");
        if (!alf_aps_ok(d.luma_aps, true) || !alf_aps_ok(d.chroma_aps, false) || !alf_aps_ok(d.cc_cb_aps, false) ||
            !alf_aps_ok(d.cc_cr_aps, false))
            return fail(ctx, RMH_INVALID_INPUT, tag + "APS out of range");
    }
    if (written_overlap(spans)) return fail(ctx, RMH_INVALID_INPUT, "a dst plane overlaps another plane of the batch");
    return RMH_OK;
}

uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

// ------------------------------------------------------------------ GPM

namespace {

rmh_status validate_gpm(rmh_ctx* ctx, const rmh_gpm_frame* frames, size_t n_frames, const rmh_gpm_pu* pus, size_t n_pu,
                        uint32_t luma_scale, uint32_t chroma_scale) {
    if ((n_frames && !frames) || (n_pu && !pus)) return fail(ctx, RMH_INVALID_INPUT, "null frame or PU array");
    if (n_frames > 65536) return fail(ctx, RMH_LIMIT_EXCEEDED, "more than 65536 frames");
    if (luma_scale > 3 || chroma_scale > 3) return fail(ctx, RMH_INVALID_INPUT, "scales must be 0..3");
    std::vector<PlaneSpan> spans;
    for (size_t f = 0; f < n_frames; f++) {
        const rmh_gpm_frame& fr = frames[f];
        const std::string tag = "frame " + std::to_string(f) + ": ";
        for (int c = 0; c < 3; c++) {
            for (const rmh_plane* q : {&fr.pred0[c], &fr.pred1[c], &fr.dst[c]}) {
                if (!plane_ok(*q) || q->bytes_per_sample != 1 || q->bit_depth != 8)
                    return fail(ctx, RMH_INVALID_INPUT, tag + "planes must be valid 8-bit planes");
                if (q->width != fr.dst[c].width || q->height != fr.dst[c].height)
                    return fail(ctx, RMH_INVALID_INPUT, tag + "prediction / output size mismatch");
            }
            add_span(spans, fr.pred0[c], false);
            add_span(spans, fr.pred1[c], false);
            add_span(spans, fr.dst[c], true);
        }
        if (fr.dst[1].width != (fr.dst[0].width + 1) / 2 || fr.dst[1].height != (fr.dst[0].height + 1) / 2 ||
            fr.dst[2].width != fr.dst[1].width || fr.dst[2].height != fr.dst[1].height)
            return fail(ctx, RMH_INVALID_INPUT, tag + "chroma planes are not 4:2:0");
    }
    if (written_overlap(spans)) return fail(ctx, RMH_INVALID_INPUT, "an output plane overlaps another plane of the batch");
    for (size_t i = 0; i < n_pu; i++) {
        const rmh_gpm_pu& p = pus[i];
        const std::string tag = "PU " + std::to_string(i) + ": ";
        if (p.log2_w < 3 || p.log2_w > 7 || p.log2_h < 3 || p.log2_h > 7)
            return fail(ctx, RMH_INVALID_INPUT, tag + "size outside 8..128");
        if (p.frame >= n_frames) return fail(ctx, RMH_INVALID_INPUT, tag + "frame index");
        if (p.x >= frames[p.frame].dst[0].width || p.y >= frames[p.frame].dst[0].height)
            return fail(ctx, RMH_INVALID_INPUT, tag + "outside the picture");
        if (p.split_dir > 63) return fail(ctx, RMH_INVALID_INPUT, tag + "split direction outside 0..63");
        if (p.reserved[0] | p.reserved[1] | p.reserved[2])
            return fail(ctx, RMH_INVALID_INPUT, tag + "reserved bytes must be zero");
    }
    return RMH_OK;
}

// derive_gpm_params_table (vvc/inter.rs:1641-1671) over GEO_ANGLE_TO_MASK
// (inter.rs:137-140): (angle, distance) of each split direction.
void gpm_params(uint8_t* angle, uint8_t* distance) {
    static const int8_t kMask[32] = {0, -1, 1, 2, 3, 4, -1, -1, 5, -1, -1, 4, 3, 2, 1, -1,
                                     0, -1, 1, 2, 3, 4, -1, -1, 5, -1, -1, 4, 3, 2, 1, -1};
    int mode = 0;
    for (int a = 0; a < 32; a++)
        for (int d = 0; d < 4; d++) {
            const int m = kMask[a];
            if ((d == 0 && a >= 16) || ((d == 2 || d == 0) && (m == 0 || m == 5)) || m == -1) continue;
            if (mode < 64) {
                angle[mode] = (uint8_t)a;
                distance[mode] = (uint8_t)d;
                mode++;
            }
        }
}

// The split line's offsets of predict_gpm_block (inter.rs:1689-1703) for a
// bw x bh block whose positions scale onto the luma grid by 1 << scale.
void gpm_offsets(int bw, int bh, int scale, int angle, int distance, int16_t* ox, int16_t* oy) {
    const int lw = bw << scale, lh = bh << scale;
    int x = -(lw >> 1), y = -(lh >> 1);
    if (distance > 0) {
        const int am = angle % 16;
        if (am == 8 || (am != 0 && lh >= lw)) {
            const int d = (distance * lh) >> 3;
            y += angle < 16 ? d : -d;
        } else {
            const int d = (distance * lw) >> 3;
            x += angle < 16 ? d : -d;
        }
    }
    *ox = (int16_t)x;
    *oy = (int16_t)y;
}

}  // namespace

RMH_EXPORT rmh_status rmh_gpm_plan_create(rmh_ctx* ctx, const rmh_gpm_frame* frames, size_t n_frames,
                                          const rmh_gpm_pu* pus, size_t n_pu, uint32_t luma_scale,
                                          uint32_t chroma_scale, rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    rmh_status v = validate_gpm(ctx, frames, n_frames, pus, n_pu, luma_scale, chroma_scale);
    if (v != RMH_OK) return v;
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    uint8_t ang[64], dis[64];
    gpm_params(ang, dis);
    std::vector<rmh::GpmItem> items;
    uint64_t bytes = 0;
    for (size_t i = 0; i < n_pu; i++) {
        const rmh_gpm_pu& pu = pus[i];
        const int w = 1 << pu.log2_w, h = 1 << pu.log2_h;
        const int a = ang[pu.split_dir], d = dis[pu.split_dir];
        rmh::GpmItem base{};
        base.px = pu.x;
        base.py = pu.y;
        base.frame = pu.frame;
        base.dx = (uint8_t)a;
        base.dy = (uint8_t)((a + 8) % 32);
        gpm_offsets(w, h, (int)luma_scale, a, d, &base.lox, &base.loy);
        gpm_offsets(w / 2, h / 2, (int)chroma_scale, a, d, &base.cox, &base.coy);
        pu_tiles(frames[pu.frame].dst, pu.x, pu.y, w, h, [&](int x0, int y0, int tw, int th, int cw, int ch) {
            rmh::GpmItem it = base;
            it.x0 = (uint16_t)x0;
            it.y0 = (uint16_t)y0;
            it.tw = (uint8_t)tw;
            it.th = (uint8_t)th;
            it.cw = (uint8_t)cw;
            it.ch = (uint8_t)ch;
            items.push_back(it);
            bytes += 3ull * (tw * th + 2 * cw * ch);  // two predictions read, one written
        });
        bytes += sizeof(rmh_gpm_pu);
    }
    if (items.size() > 0x7fffffffu) return fail(ctx, RMH_LIMIT_EXCEEDED, "too many GPM work items");
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kGpm;
    p->n_gpm_items = items.size();
    p->gpm_scale[0] = luma_scale;
    p->gpm_scale[1] = chroma_scale;
    p->bytes = bytes;
    Blob blob;
    const size_t o_frames = blob.add(frames, n_frames * sizeof(rmh_gpm_frame));
    const size_t o_items = blob.add(items.data(), items.size() * sizeof(rmh::GpmItem));
    std::string err;
    if (!upload_blob(p, blob, &err)) {
        delete p;
        return fail(ctx, RMH_EXECUTION, "gpm plan upload failed: " + err);
    }
    p->d_gpm_frames = (const rmh_gpm_frame*)((char*)p->d_mem + o_frames);
    p->d_gpm_items = (const rmh::GpmItem*)((char*)p->d_mem + o_items);
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_gpm_plan_run(rmh_plan* plan, void* stream) {
    if (!plan || plan->kind != kGpm) return RMH_INVALID_INPUT;
    if (plan->n_gpm_items == 0) return RMH_OK;
    return run_plan(plan, nullptr, nullptr, 0, 0, stream);
}

RMH_EXPORT rmh_status rmh_lmcs_plan_create(rmh_ctx* ctx, const rmh_lmcs_frame* frames, size_t n_frames,
                                           rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    rmh_status v = validate_lmcs(ctx, frames, n_frames);
    if (v != RMH_OK) return v;
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kLmcs;
    // jobs of 8-bit planes first; LUTs after the job and workgroup tables
    std::vector<size_t> order;
    for (int pass = 1; pass <= 2; pass++)
        for (size_t f = 0; f < n_frames; f++)
            if (frames[f].src.bytes_per_sample == pass) order.push_back(f);
    std::vector<rmh::LmcsJob> jobs;
    std::vector<uint32_t> block_job;
    std::vector<uint16_t> luts;
    std::vector<uint32_t> lut_at;
    uint32_t blocks8 = 0;
    for (size_t f : order) {
        const rmh_lmcs_frame& d = frames[f];
        const int bd = d.src.bit_depth;
        rmh::LmcsJob j{};
        j.src = d.src.data;
        j.dst = d.dst.data;
        j.width = d.src.width;
        j.height = d.src.height;
        j.src_stride = d.src.stride;
        j.dst_stride = d.dst.stride;
        j.bps = d.src.bytes_per_sample;
        j.cap = (1u << bd) - 1;
        j.first_block = (uint32_t)block_job.size();
        lut_at.push_back((uint32_t)luts.size());
        luts.resize(luts.size() + j.cap + 1);
        uint16_t* lut = luts.data() + lut_at.back();
        rmh::lmcs_lut(d.aps, bd, d.direction == RMH_LMCS_FORWARD ? lut : nullptr,
                      d.direction == RMH_LMCS_INVERSE ? lut : nullptr, nullptr, nullptr);
        const uint32_t nb = div_up(d.src.height, rmh::kLmcsRows);
        block_job.insert(block_job.end(), nb, (uint32_t)jobs.size());
        if (d.src.bytes_per_sample == 1) blocks8 += nb;
        jobs.push_back(j);
        p->bytes += 2 * plane_bytes(d.src);
    }
    p->lmcs_range[0] = 0;
    p->lmcs_range[1] = blocks8;
    p->lmcs_range[2] = (uint32_t)block_job.size();
    Blob blob;
    const size_t o_jobs = blob.add(jobs.data(), jobs.size() * sizeof(rmh::LmcsJob));
    const size_t o_bj = blob.add(block_job.data(), block_job.size() * 4);
    const size_t o_lut = blob.add(luts.data(), luts.size() * 2);
    // the jobs' LUT addresses are known once the device blob exists: patch and re-upload
    std::string err;
    if (!upload_blob(p, blob, &err)) {
        delete p;
        return fail(ctx, RMH_EXECUTION, "lmcs plan upload failed: " + err);
    }
    if (!jobs.empty()) {
        char* base = (char*)p->d_mem;
        for (size_t k = 0; k < jobs.size(); k++) jobs[k].lut = (const uint16_t*)(base + o_lut) + lut_at[k];
        hipError_t e = hipMemcpy(base + o_jobs, jobs.data(), jobs.size() * sizeof(rmh::LmcsJob), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(p->d_mem);
            delete p;
            return fail(ctx, RMH_EXECUTION, std::string("lmcs plan upload failed: ") + hipGetErrorString(e));
        }
        p->d_lmcs_jobs = (const rmh::LmcsJob*)(base + o_jobs);
        p->d_lmcs_block_job = (const uint32_t*)(base + o_bj);
    }
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_lmcs_plan_run(rmh_plan* plan, void* stream) {
    if (!plan || plan->kind != kLmcs) return RMH_INVALID_INPUT;
    if (plan->lmcs_range[2] == 0) return RMH_OK;
    return run_plan(plan, nullptr, nullptr, 0, 0, stream);
}

RMH_EXPORT rmh_status rmh_alf_plan_create(rmh_ctx* ctx, const rmh_alf_frame* frames, size_t n_frames, rmh_plan** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    rmh_status v = validate_alf(ctx, frames, n_frames);
    if (v != RMH_OK) return v;
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    rmh_plan* p = new rmh_plan();
    p->ctx = ctx;
    p->kind = kAlf;
    std::vector<size_t> order;
    for (int pass = 1; pass <= 2; pass++)
        for (size_t f = 0; f < n_frames; f++)
            if (frames[f].src[0].bytes_per_sample == pass) order.push_back(f);
    // per frame (in job order): luma bank, chroma bank, CCALF banks (Cb, Cr)
    std::vector<int16_t> banks;
    struct BankAt {
        size_t luma, chroma, cc[2];
    };
    std::vector<BankAt> at;
    std::vector<rmh::AlfLumaJob> lj;
    std::vector<rmh::AlfChromaJob> cj;
    std::vector<uint32_t> lbj, cbj;
    uint32_t l8 = 0, c8 = 0;
    for (size_t f : order) {
        const rmh_alf_frame& d = frames[f];
        const int bd = d.src[0].bit_depth;
        const uint32_t ctb_log2 = d.ctb_size == 16 ? 4 : d.ctb_size == 32 ? 5 : d.ctb_size == 64 ? 6 : 7;
        BankAt b{};
        b.luma = banks.size();
        banks.resize(banks.size() + rmh::kAlfBankEntries);
        rmh::alf_luma_bank(d.luma_aps, bd, banks.data() + b.luma);
        b.chroma = banks.size();
        banks.resize(banks.size() + 8 * 12);
        if (d.chroma_aps) rmh::alf_chroma_bank(*d.chroma_aps, bd, banks.data() + b.chroma);
        for (int c = 0; c < 2; c++) {
            const rmh_alf_aps* a = c == 0 ? d.cc_cb_aps : d.cc_cr_aps;
            b.cc[c] = banks.size();
            banks.resize(banks.size() + 4 * 7);
            if (a) memcpy(banks.data() + b.cc[c], c == 0 ? &a->cc_cb[0][0] : &a->cc_cr[0][0], 4 * 7 * sizeof(int16_t));
        }
        at.push_back(b);
        const rmh_plane& y = d.src[0];
        rmh::AlfLumaJob L{};
        L.src = y.data;
        L.dst = d.dst[0].data;
        L.ctbs = d.ctbs;
        L.src_stride = y.stride;
        L.dst_stride = d.dst[0].stride;
        L.width = y.width;
        L.height = y.height;
        L.vis_w = d.vis_w;
        L.vis_h = d.vis_h;
        L.bw = std::min(d.vis_w, y.width);
        L.bh = std::min(d.vis_h, y.height);
        L.bps = y.bytes_per_sample;
        L.bd = (uint32_t)bd;
        L.ctb_log2 = ctb_log2;
        L.ctb_cols = d.ctb_cols;
        L.ctb_xn = div_up(d.vis_w, d.ctb_size);
        L.ctb_yn = div_up(d.vis_h, d.ctb_size);
        L.has_aps = d.luma_aps != nullptr;
        L.tiles_x = div_up(y.width, rmh::kAlfTileW);
        L.first_block = (uint32_t)lbj.size();
        const uint32_t nl = L.tiles_x * div_up(y.height, rmh::kAlfTileH);
        lbj.insert(lbj.end(), nl, (uint32_t)lj.size());
        if (y.bytes_per_sample == 1) l8 += nl;
        lj.push_back(L);
        for (int c = 1; c < 3; c++) {
            const rmh_plane& pc = d.src[c];
            const rmh_alf_aps* cc = c == 1 ? d.cc_cb_aps : d.cc_cr_aps;
            rmh::AlfChromaJob C{};
            C.src = pc.data;
            C.dst = d.dst[c].data;
            C.luma = y.data;
            C.ctbs = d.ctbs;
            C.src_stride = pc.stride;
            C.dst_stride = d.dst[c].stride;
            C.luma_stride = y.stride;
            C.width = pc.width;
            C.height = pc.height;
            C.cvis_w = (d.vis_w + 1) / 2;
            C.cvis_h = (d.vis_h + 1) / 2;
            C.cbw = std::min(C.cvis_w, pc.width);
            C.cbh = std::min(C.cvis_h, pc.height);
            C.lw = L.bw;
            C.lh = L.bh;
            C.vis_w = d.vis_w;
            C.vis_h = d.vis_h;
            C.bps = pc.bytes_per_sample;
            C.bd = (uint32_t)bd;
            C.ctb_log2 = ctb_log2;
            C.ctb_cols = d.ctb_cols;
            C.ctb_xn = L.ctb_xn;
            C.ctb_yn = L.ctb_yn;
            C.comp = (uint32_t)c;
            C.n_chroma = d.chroma_aps ? d.chroma_aps->n_chroma : 0;
            C.n_cc = cc ? (c == 1 ? cc->n_cc_cb : cc->n_cc_cr) : 0;
            C.tiles_x = div_up(pc.width, rmh::kAlfTileW);
            C.first_block = (uint32_t)cbj.size();
            const uint32_t nc = C.tiles_x * div_up(pc.height, rmh::kAlfTileH);
            cbj.insert(cbj.end(), nc, (uint32_t)cj.size());
            if (pc.bytes_per_sample == 1) c8 += nc;
            cj.push_back(C);
        }
        for (int c = 0; c < 3; c++) p->bytes += 2 * plane_bytes(d.src[c]);
    }
    p->alf_luma_range[1] = l8;
    p->alf_luma_range[2] = (uint32_t)lbj.size();
    p->alf_chroma_range[1] = c8;
    p->alf_chroma_range[2] = (uint32_t)cbj.size();
    Blob blob;
    const size_t o_lj = blob.add(lj.data(), lj.size() * sizeof(rmh::AlfLumaJob));
    const size_t o_cj = blob.add(cj.data(), cj.size() * sizeof(rmh::AlfChromaJob));
    const size_t o_lbj = blob.add(lbj.data(), lbj.size() * 4);
    const size_t o_cbj = blob.add(cbj.data(), cbj.size() * 4);
    const size_t o_bank = blob.add(banks.data(), banks.size() * 2);
    std::string err;
    if (!upload_blob(p, blob, &err)) {
        delete p;
        return fail(ctx, RMH_EXECUTION, "alf plan upload failed: " + err);
    }
    if (!lj.empty()) {
        char* base = (char*)p->d_mem;
        const int16_t* bank = (const int16_t*)(base + o_bank);
        for (size_t k = 0; k < lj.size(); k++) {
            const rmh_alf_frame& d = frames[order[k]];
            lj[k].bank = bank + at[k].luma;
            for (int c = 0; c < 2; c++) {
                rmh::AlfChromaJob& C = cj[2 * k + c];
                C.chroma_bank = d.chroma_aps && d.chroma_aps->n_chroma ? bank + at[k].chroma : nullptr;
                C.cc_bank = C.n_cc ? bank + at[k].cc[c] : nullptr;
            }
        }
        hipError_t e = hipMemcpy(base + o_lj, lj.data(), lj.size() * sizeof(rmh::AlfLumaJob), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(base + o_cj, cj.data(), cj.size() * sizeof(rmh::AlfChromaJob), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(p->d_mem);
            delete p;
            return fail(ctx, RMH_EXECUTION, std::string("alf plan upload failed: ") + hipGetErrorString(e));
        }
        p->d_alf_luma = (const rmh::AlfLumaJob*)(base + o_lj);
        p->d_alf_chroma = (const rmh::AlfChromaJob*)(base + o_cj);
        p->d_alf_luma_bj = (const uint32_t*)(base + o_lbj);
        p->d_alf_chroma_bj = (const uint32_t*)(base + o_cbj);
    }
    *out = p;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_alf_plan_run(rmh_plan* plan, void* stream) {
    if (!plan || plan->kind != kAlf) return RMH_INVALID_INPUT;
    if (plan->alf_luma_range[2] == 0) return RMH_OK;
    return run_plan(plan, nullptr, nullptr, 0, 0, stream);
}

// ------------------------------------------------------------ plans, graphs

RMH_EXPORT void rmh_plan_destroy(rmh_plan* plan) {
    if (!plan) return;
    DeviceGuard dg(plan->ctx->device);
    if (plan->d_mem) hipFree(plan->d_mem);
    if (plan->inverse) rmh_plan_destroy(plan->inverse);
    delete plan;
}

RMH_EXPORT uint64_t rmh_plan_bytes(const rmh_plan* plan) { return plan ? plan->bytes : 0; }

RMH_EXPORT rmh_status rmh_graph_create(rmh_ctx* ctx, const rmh_graph_node* nodes, size_t n_nodes, rmh_graph** out) {
    if (!ctx || !out || (n_nodes && !nodes)) return RMH_INVALID_INPUT;
    *out = nullptr;
    for (size_t i = 0; i < n_nodes; i++) {
        if (!nodes[i].plan || nodes[i].plan->ctx != ctx) return fail(ctx, RMH_INVALID_INPUT, "graph node without a plan of this context");
        if ((nodes[i].plan->kind == kResidual || nodes[i].plan->kind == kForward) && nodes[i].plan->n_tu &&
            (!nodes[i].d_coeff || (!nodes[i].d_residual && !nodes[i].plan->d_place)))
            return fail(ctx, RMH_INVALID_INPUT, "residual graph node without device buffers");
    }
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    DeviceGuard dg(ctx->device);
    rmh_graph* g = new rmh_graph();
    g->ctx = ctx;
    hipError_t e = hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking);
    auto event = [&]() {
        hipEvent_t ev = nullptr;
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (ev) g->events.push_back(ev);
        return ev;
    };
    if (e == hipSuccess) e = hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal);
    std::string err;
    rmh_status st = RMH_OK;
    if (e == hipSuccess) {
        auto empty = [](const rmh_plan* p) {
            return ((p->kind == kResidual || p->kind == kForward) && p->n_tu == 0) ||
                   (p->kind == kMc && p->n_items == 0 && p->n_ref_items == 0) || (p->kind == kGpm && p->n_gpm_items == 0) ||
                   (p->kind == kDeblock && p->n_dbk == 0) || (p->kind == kSao && p->n_sao_jobs == 0) ||
                   (p->kind == kEdge && p->total_cells == 0) || (p->kind == kLmcs && p->lmcs_range[2] == 0) ||
                   (p->kind == kAlf && p->alf_luma_range[2] == 0);
        };
        // Residual plans run in parts on two branches of the step: the
        // prepass depends only on the coefficients, so it starts with the
        // step, beside the nodes before it (MC); the generic kernel's TUs and
        // the tiles' TUs are disjoint, so the generic kernel runs beside the
        // tiles. Each branch joins the main one before the node that follows.
        std::vector<hipEvent_t> prep_done(n_nodes, nullptr);
        hipEvent_t start = event();
        if (e == hipSuccess) e = hipEventRecord(start, g->capture);
        if (e == hipSuccess) e = hipStreamWaitEvent(g->side, start, 0);
        bool side_used = false;
        for (size_t i = 0; i < n_nodes && st == RMH_OK && e == hipSuccess; i++) {
            rmh_plan* p = nodes[i].plan;
            if (empty(p) || !residual_tiled_run(p, nodes[i].d_coeff, nodes[i].d_residual) || !p->n_prep) continue;
            st = run_plan_locked_free(p, nodes[i].d_coeff, nodes[i].d_residual, 0, p->n_tu, g->side, &err, kPartPrepass);
            prep_done[i] = event();
            if (e == hipSuccess) e = hipEventRecord(prep_done[i], g->side);
            side_used = true;
        }
        for (size_t i = 0; i < n_nodes && st == RMH_OK && e == hipSuccess; i++) {
            rmh_plan* p = nodes[i].plan;
            if (empty(p)) continue;
            if (!residual_tiled_run(p, nodes[i].d_coeff, nodes[i].d_residual)) {
                st = run_plan_locked_free(p, nodes[i].d_coeff, nodes[i].d_residual, 0, p->n_tu, g->capture, &err);
                continue;
            }
            if (prep_done[i]) e = hipStreamWaitEvent(g->capture, prep_done[i], 0);
            const bool generic = p->n_tu > p->n_fast;
            if (generic && e == hipSuccess) {
                hipEvent_t fork = event();
                if (e == hipSuccess) e = hipEventRecord(fork, g->capture);
                if (e == hipSuccess) e = hipStreamWaitEvent(g->side, fork, 0);
                if (e == hipSuccess)
                    st = run_plan_locked_free(p, nodes[i].d_coeff, nodes[i].d_residual, 0, p->n_tu, g->side, &err,
                                              kPartGeneric);
                side_used = true;
            }
            if (e == hipSuccess && st == RMH_OK)
                st = run_plan_locked_free(p, nodes[i].d_coeff, nodes[i].d_residual, 0, p->n_tu, g->capture, &err,
                                          kPartTiles);
            if (generic && e == hipSuccess && st == RMH_OK) {
                hipEvent_t join = event();
                if (e == hipSuccess) e = hipEventRecord(join, g->side);
                if (e == hipSuccess) e = hipStreamWaitEvent(g->capture, join, 0);
            }
        }
        // the side branch always rejoins before the capture ends
        if (e == hipSuccess) {
            hipEvent_t end = event();
            if (e == hipSuccess) e = hipEventRecord(end, g->side);
            if (e == hipSuccess) e = hipStreamWaitEvent(g->capture, end, 0);
        }
        (void)side_used;
        hipError_t e2 = hipStreamEndCapture(g->capture, &g->graph);
        if (e2 != hipSuccess && st == RMH_OK && e == hipSuccess) e = e2;
    }
    if (e == hipSuccess && st == RMH_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess || st != RMH_OK) {
        if (g->exec) hipGraphExecDestroy(g->exec);
        if (g->graph) hipGraphDestroy(g->graph);
        if (g->capture) hipStreamDestroy(g->capture);
        if (g->side) hipStreamDestroy(g->side);
        for (hipEvent_t ev : g->events) hipEventDestroy(ev);
        delete g;
        return fail(ctx, RMH_EXECUTION, st != RMH_OK ? err : std::string("graph capture failed: ") + hipGetErrorString(e));
    }
    for (size_t i = 0; i < n_nodes; i++) {
        g->plans.push_back(nodes[i].plan);
        nodes[i].plan->graph_refs.fetch_add(1);
    }
    *out = g;
    return RMH_OK;
}

RMH_EXPORT rmh_status rmh_graph_launch(rmh_graph* g, void* stream) {
    if (!g) return RMH_INVALID_INPUT;
    rmh_status gt = gate(g->ctx);
    if (gt != RMH_OK) return fail(g->ctx, gt, gt == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(g->ctx->device);
    hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
    if (e != hipSuccess) return fail(g->ctx, RMH_EXECUTION, std::string("graph launch failed: ") + hipGetErrorString(e));
    return RMH_OK;
}

RMH_EXPORT void rmh_graph_destroy(rmh_graph* g) {
    if (!g) return;
    DeviceGuard dg(g->ctx->device);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->capture) hipStreamDestroy(g->capture);
    if (g->side) hipStreamDestroy(g->side);
    for (hipEvent_t ev : g->events) hipEventDestroy(ev);
    for (rmh_plan* p : g->plans) p->graph_refs.fetch_sub(1);
    delete g;
}

// ------------------------------------------------------------ CU-split CNN

struct rmh_cnn {
    rmh_ctx* ctx = nullptr;
    uint32_t max_batch = 0;
    void* d_mem = nullptr;
    rmh::cnn::Buffers b{};
    uint8_t* d_blocks = nullptr;  // staging of rmh_cnn_predict
    float* d_qp = nullptr;
    float* d_probs = nullptr;
    float* h_probs = nullptr;  // pinned read-back of rmh_cnn_predict (max_batch * 6 floats)
    int step = 0;
    // rmh_cnn_predict runs on the network's own stream under its own lock,
    // so predictions of different instances overlap on the device
    hipStream_t stream = nullptr;
    std::mutex mu;
};

namespace {

// splitmix64 stream -> standard normal (Box-Muller), for seeded He initialisation.
struct InitRng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

void init_params(uint64_t seed, std::vector<float>& p) {
    using namespace rmh::cnn;
    p.assign(kParams, 0.0f);
    InitRng r{seed};
    auto fill = [&](int off, int count, double fan_in) {
        const double sd = std::sqrt(2.0 / fan_in);
        for (int i = 0; i < count; i++) p[off + i] = (float)(sd * r.normal());
    };
    fill(kOffW1, 32 * 9, 9.0);
    fill(kOffW2, 32 * 9 * 32, 288.0);
    fill(kOffW3, 32 * 9 * 32, 288.0);
    fill(kOffW4, 64 * 33, 33.0);
    fill(kOffW5, 6 * 64, 128.0);
}

rmh_status cnn_launch_status(rmh_cnn* c, hipError_t e, const char* what) {
    if (e != hipSuccess) return fail(c->ctx, RMH_EXECUTION, std::string(what) + ": " + hipGetErrorString(e));
    return RMH_OK;
}

}  // namespace

RMH_EXPORT size_t rmh_cnn_param_count(void) { return (size_t)rmh::cnn::kParams; }

RMH_EXPORT rmh_status rmh_cnn_create(rmh_ctx* ctx, uint32_t max_batch, uint64_t seed, rmh_cnn** out) {
    if (!ctx || !out) return RMH_INVALID_INPUT;
    *out = nullptr;
    if (max_batch == 0 || max_batch > 65536) return fail(ctx, RMH_LIMIT_EXCEEDED, "max_batch must be 1..65536");
    if (ctx->disabled.load()) return fail(ctx, RMH_UNAVAILABLE, "accelerator disabled");
    using namespace rmh::cnn;
    DeviceGuard dg(ctx->device);
    const size_t n = max_batch, act = n * 1024 * 32 * 2;
    const size_t sizes[] = {
        kParams * 4u, kParams * 4u, kParams * 4u, kParams * 4u,  // params, grads, m, v
        32 * 16 * 2, 9216 * 2, 9216 * 2, 9216 * 2, 9216 * 2,    // w1 w2 w3 w2t w3t
        n * 1024 * 2, act, act, act, act, act,                  // x a1 a2 a3 dz dz2
        n * 32 * 4, n * 32 * 4, n * 6 * 4, n * 9216 * 4,        // pooled dpool probs part
        n * 1024, n * 4, n * 6 * 4};                            // staging: blocks qp probs
    size_t total = 0;
    for (size_t s : sizes) total += aligned(s);
    rmh_cnn* c = new rmh_cnn();
    c->ctx = ctx;
    c->max_batch = max_batch;
    hipError_t e = hipMalloc(&c->d_mem, total);
    if (e == hipSuccess) {
        e = hipHostMalloc((void**)&c->h_probs, n * 6 * sizeof(float), hipHostMallocDefault);
        if (e != hipSuccess) {
            hipFree(c->d_mem);
            c->h_probs = nullptr;
        }
    }
    if (e == hipSuccess) {
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            hipFree(c->d_mem);
            hipHostFree(c->h_probs);
            c->stream = nullptr;
        }
    }
    if (e != hipSuccess) {
        delete c;
        return fail(ctx, RMH_LIMIT_EXCEEDED, std::string("device allocation: ") + hipGetErrorString(e));
    }
    char* p = (char*)c->d_mem;
    void* ptrs[sizeof sizes / sizeof sizes[0]];
    for (size_t i = 0; i < sizeof sizes / sizeof sizes[0]; i++) {
        ptrs[i] = p;
        p += aligned(sizes[i]);
    }
    Buffers& b = c->b;
    b.params = (float*)ptrs[0];
    b.grads = (float*)ptrs[1];
    b.adam_m = (float*)ptrs[2];
    b.adam_v = (float*)ptrs[3];
    b.w1 = (uint16_t*)ptrs[4];
    b.w2 = (uint16_t*)ptrs[5];
    b.w3 = (uint16_t*)ptrs[6];
    b.w2t = (uint16_t*)ptrs[7];
    b.w3t = (uint16_t*)ptrs[8];
    b.x = (uint16_t*)ptrs[9];
    b.a1 = (uint16_t*)ptrs[10];
    b.a2 = (uint16_t*)ptrs[11];
    b.a3 = (uint16_t*)ptrs[12];
    b.dz = (uint16_t*)ptrs[13];
    b.dz2 = (uint16_t*)ptrs[14];
    b.pooled = (float*)ptrs[15];
    b.dpool = (float*)ptrs[16];
    b.probs = (float*)ptrs[17];
    b.part = (float*)ptrs[18];
    c->d_blocks = (uint8_t*)ptrs[19];
    c->d_qp = (float*)ptrs[20];
    c->d_probs = (float*)ptrs[21];
    std::vector<float> init;
    init_params(seed, init);
    e = hipMemcpy(b.params, init.data(), kParams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b.adam_m, 0, kParams * 4);
    if (e == hipSuccess) e = hipMemset(b.adam_v, 0, kParams * 4);
    if (e == hipSuccess) e = pack_weights(b, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        hipStreamDestroy(c->stream);
        hipHostFree(c->h_probs);
        hipFree(c->d_mem);
        delete c;
        return fail(ctx, RMH_EXECUTION, std::string("network initialisation: ") + hipGetErrorString(e));
    }
    *out = c;
    return RMH_OK;
}

RMH_EXPORT void rmh_cnn_destroy(rmh_cnn* c) {
    if (!c) return;
    DeviceGuard dg(c->ctx->device);
    if (c->stream) {
        hipStreamSynchronize(c->stream);
        hipStreamDestroy(c->stream);
    }
    if (c->h_probs) hipHostFree(c->h_probs);
    if (c->d_mem) hipFree(c->d_mem);
    delete c;
}

RMH_EXPORT rmh_status rmh_cnn_get_params(rmh_cnn* c, float* out) {
    if (!c || !out) return RMH_INVALID_INPUT;
    DeviceGuard dg(c->ctx->device);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, c->b.params, rmh::cnn::kParams * 4, hipMemcpyDeviceToHost);
    return cnn_launch_status(c, e, "parameter download");
}

RMH_EXPORT rmh_status rmh_cnn_set_params(rmh_cnn* c, const float* in) {
    if (!c || !in) return RMH_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->ctx->device);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(c->b.params, in, rmh::cnn::kParams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(c->b.adam_m, 0, rmh::cnn::kParams * 4);
    if (e == hipSuccess) e = hipMemset(c->b.adam_v, 0, rmh::cnn::kParams * 4);
    if (e == hipSuccess) e = rmh::cnn::pack_weights(c->b, c->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->ctx->stream);
    c->step = 0;
    return cnn_launch_status(c, e, "parameter upload");
}

RMH_EXPORT rmh_status rmh_cnn_forward(rmh_cnn* c, const uint8_t* d_blocks, const float* d_qp, uint32_t n,
                                      float* d_probs, void* stream) {
    if (!c) return RMH_INVALID_INPUT;
    if (n == 0) return RMH_OK;
    if (!d_blocks || !d_qp || !d_probs) return fail(c->ctx, RMH_INVALID_INPUT, "null device buffer");
    if (n > c->max_batch) return fail(c->ctx, RMH_LIMIT_EXCEEDED, "batch larger than max_batch");
    // the activation buffers are the network's: one call at a time enqueues on them
    std::lock_guard<std::mutex> lk(c->mu);
    rmh_status g = gate(c->ctx);
    if (g != RMH_OK) return fail(c->ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(c->ctx->device);
    return cnn_launch_status(c, rmh::cnn::forward(c->b, d_blocks, d_qp, (int)n, d_probs, (hipStream_t)stream),
                             "forward launch");
}

RMH_EXPORT rmh_status rmh_cnn_predict(rmh_cnn* c, const uint8_t* blocks, const float* qp, uint32_t n, float* probs) {
    if (!c) return RMH_INVALID_INPUT;
    if (n == 0) return RMH_OK;
    if (!blocks || !qp || !probs) return fail(c->ctx, RMH_INVALID_INPUT, "null host buffer");
    if (n > c->max_batch) return fail(c->ctx, RMH_LIMIT_EXCEEDED, "batch larger than max_batch");
    rmh_ctx* ctx = c->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    rmh_status g = gate(ctx);
    if (g != RMH_OK) return fail(ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(ctx->device);
    hipStream_t s = c->stream;
    hipError_t e = hipMemcpyAsync(c->d_blocks, blocks, (size_t)n * 1024, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_qp, qp, (size_t)n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = rmh::cnn::forward(c->b, c->d_blocks, c->d_qp, (int)n, c->d_probs, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->h_probs, c->d_probs, (size_t)n * 24, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return fail(ctx, RMH_EXECUTION, std::string("predict: ") + hipGetErrorString(e));
    // the context's timeout, without holding the context lock
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(ctx->timeout_ms);
    for (;;) {
        e = hipStreamQuery(s);
        if (e == hipSuccess) {
            // the caller's buffer is written only once the copy has landed
            memcpy(probs, c->h_probs, (size_t)n * 24);
            return RMH_OK;
        }
        if (e != hipErrorNotReady) {
            ctx->disabled.store(true);
            return fail(ctx, RMH_EXECUTION, std::string("device execution failed: ") + hipGetErrorString(e));
        }
        if (std::chrono::steady_clock::now() > deadline) {
            ctx->disabled.store(true);
            return fail(ctx, RMH_TIMEOUT, "device work did not finish within the timeout");
        }
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

namespace {

rmh_status cnn_gradients(rmh_cnn* c, const uint8_t* d_blocks, const float* d_qp, const uint8_t* d_labels,
                         const float* d_targets, uint32_t n, float* d_grads, float* d_loss, void* stream) {
    if (!c) return RMH_INVALID_INPUT;
    if (n == 0) return fail(c->ctx, RMH_INVALID_INPUT, "empty batch");
    if (!d_blocks || !d_qp || !(d_labels || d_targets) || !d_grads || !d_loss)
        return fail(c->ctx, RMH_INVALID_INPUT, "null device buffer");
    if (n > c->max_batch) return fail(c->ctx, RMH_LIMIT_EXCEEDED, "batch larger than max_batch");
    rmh_status g = gate(c->ctx);
    if (g != RMH_OK) return fail(c->ctx, g, g == RMH_UNAVAILABLE ? "accelerator disabled" : "injected device failure");
    DeviceGuard dg(c->ctx->device);
    return cnn_launch_status(c,
                             rmh::cnn::gradients(c->b, d_blocks, d_qp, d_labels, d_targets, (int)n, d_grads, d_loss,
                                                 (hipStream_t)stream),
                             "gradient launch");
}

rmh_status cnn_step(rmh_cnn* c, float lr, void* stream) {
    DeviceGuard dg(c->ctx->device);
    c->step++;
    return cnn_launch_status(c, rmh::cnn::adam(c->b, c->step, lr, (hipStream_t)stream), "optimiser launch");
}

}  // namespace

RMH_EXPORT rmh_status rmh_cnn_gradients(rmh_cnn* c, const uint8_t* d_blocks, const float* d_qp, const uint8_t* d_labels,
                                        uint32_t n, float* d_grads, float* d_loss, void* stream) {
    if (!c) return RMH_INVALID_INPUT;
    if (!d_labels) return fail(c->ctx, RMH_INVALID_INPUT, "null labels");
    std::lock_guard<std::mutex> lk(c->mu);
    return cnn_gradients(c, d_blocks, d_qp, d_labels, nullptr, n, d_grads, d_loss, stream);
}

RMH_EXPORT rmh_status rmh_cnn_train_step(rmh_cnn* c, const uint8_t* d_blocks, const float* d_qp,
                                         const uint8_t* d_labels, uint32_t n, float lr, float* d_loss, void* stream) {
    if (!c) return RMH_INVALID_INPUT;
    if (!(lr > 0.0f) || lr > 1.0f) return fail(c->ctx, RMH_INVALID_INPUT, "learning rate outside (0, 1]");
    if (!d_labels) return fail(c->ctx, RMH_INVALID_INPUT, "null labels");
    std::lock_guard<std::mutex> lk(c->mu);
    rmh_status st = cnn_gradients(c, d_blocks, d_qp, d_labels, nullptr, n, c->b.grads, d_loss, stream);
    return st != RMH_OK ? st : cnn_step(c, lr, stream);
}

RMH_EXPORT rmh_status rmh_cnn_train_step_soft(rmh_cnn* c, const uint8_t* d_blocks, const float* d_qp,
                                              const float* d_targets, uint32_t n, float lr, float* d_loss,
                                              void* stream) {
    if (!c) return RMH_INVALID_INPUT;
    if (!(lr > 0.0f) || lr > 1.0f) return fail(c->ctx, RMH_INVALID_INPUT, "learning rate outside (0, 1]");
    if (!d_targets) return fail(c->ctx, RMH_INVALID_INPUT, "null targets");
    std::lock_guard<std::mutex> lk(c->mu);
    rmh_status st = cnn_gradients(c, d_blocks, d_qp, nullptr, d_targets, n, c->b.grads, d_loss, stream);
    return st != RMH_OK ? st : cnn_step(c, lr, stream);
}
