// Address models of the library's kernels for the served-path replay (test infrastructure).
//
// Each launcher of recon_internal.h is replaced by a stream operation that, when it runs, reads the
// launch's job records and tile / band tables from "device" memory exactly where the kernel reads
// them, follows the kernel's own index math (the file and the expression are named at each check)
// and checks every buffer range the kernel would read or write against the allocations that are
// live at that moment. Data-dependent ranges (the largest level record a picture's macroblocks
// name, the reference pictures its blocks use) come from scanning the records the kernel would
// read; the scans are cached per buffer content. It also checks that every launch covers each of
// its jobs exactly once (tile tables, row ranges, band layouts), since a wrong table sends a
// workgroup to the wrong job or past a job's last block.
#include <algorithm>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <vector>

#include "model.h"
#include "../../mmrecode_amd/csrc/recon_internal.h"

using model::check;
using model::violation;

namespace {

template <typename T>
bool read_rec(const T* p, T& out, const char* what) {
    if (!check(p, sizeof(T), what)) return false;
    memcpy(&out, p, sizeof(T));
    return true;
}

bool group_blocks_ok(uint32_t g) { return g == 16 || g == 32 || g == 64; }

// The tile offsets a launch gave job j must be exactly 0 .. n - 1.
void check_tiles(const char* kernel, uint32_t j, std::vector<uint32_t>& offs, uint64_t n) {
    std::sort(offs.begin(), offs.end());
    bool ok = offs.size() == n;
    for (size_t i = 0; ok && i < offs.size(); ++i) ok = offs[i] == i;
    if (!ok)
        violation("%s: job %u got %zu tiles (first offset %u), needs offsets 0..%llu once each", kernel, j, offs.size(),
                  offs.empty() ? 0u : offs[0], (unsigned long long)n - 1);
}

// Tiles of a block kernel: job index -> tile offsets (tile - job.tile_begin).
template <typename Job>
bool walk_tiles(const char* kernel, const Job* jobs, const uint32_t* tile_job, uint32_t tile_base, uint32_t tiles,
                std::map<uint32_t, std::vector<uint32_t>>& offs, std::map<uint32_t, Job>& recs) {
    for (uint64_t t = tile_base; t < (uint64_t)tile_base + tiles; ++t) {
        uint32_t j;
        if (!read_rec(tile_job + t, j, kernel)) return false;
        auto it = recs.find(j);
        if (it == recs.end()) {
            Job job;
            if (!read_rec(jobs + j, job, kernel)) return false;
            it = recs.emplace(j, job).first;
        }
        offs[j].push_back((uint32_t)(t - it->second.tile_begin));
    }
    return true;
}

uint32_t popc(uint32_t v) { return (uint32_t)__builtin_popcount(v); }

// ---- MPEG-2 macroblock scan (mpeg2.hip: mpeg2_order_kernel's entries, mpeg2_kernel's uses)
struct Mpeg2Scan {
    int64_t max_level = -1;    // largest level record index of a coded block (24-bit field, as the kernel keeps it)
    bool fwd = false, bwd = false, field = false, bad_index = false;
};
std::map<std::tuple<const void*, uint64_t, uint32_t>, Mpeg2Scan> g_mpeg2_scans;

Mpeg2Scan scan_mpeg2(const mmr_recon_mpeg2_macroblock* mbs, uint32_t n) {
    const auto key = std::make_tuple((const void*)mbs, model::content_key(mbs), n);
    auto it = g_mpeg2_scans.find(key);
    if (it != g_mpeg2_scans.end()) return it->second;
    Mpeg2Scan r;
    for (uint32_t i = 0; i < n; ++i) {
        const mmr_recon_mpeg2_macroblock& m = mbs[i];
        const bool intra = m.flags & MMR_MPEG2_MB_INTRA;
        const uint32_t cbp = intra ? 0x3fu : (m.cbp & 0x3fu);
        for (uint32_t b = 0; b < 6; ++b) {
            if (!((cbp >> (5 - b)) & 1u)) continue;
            const uint32_t index = m.block_offset + (intra ? b : popc((m.cbp & 0x3fu) >> (6 - b)));
            if (index >= (1u << 24)) r.bad_index = true;   // reported by the order pass; the main kernel keeps 24 bits
            r.max_level = std::max<int64_t>(r.max_level, index & 0xffffffu);
        }
        if (!intra) {
            r.fwd |= (m.flags & MMR_MPEG2_MB_FORWARD) != 0;
            r.bwd |= (m.flags & MMR_MPEG2_MB_BACKWARD) != 0;
            r.field |= (m.flags & MMR_MPEG2_MB_FIELD_MOTION) != 0;
        }
    }
    if (g_mpeg2_scans.size() > 4096) g_mpeg2_scans.clear();
    g_mpeg2_scans.emplace(key, r);
    return r;
}

// ---- H.264 macroblock / block scan (h264.hip: h264_mb_kernel)
struct H264Scan {
    int64_t max_coef = -1, max_coef_intra = -1;   // residual macroblocks: inter ones / intra ones
    bool intra = false;
    // reference index pairs (r0 + 1, r1 + 1) of the 4x4 blocks of non-intra / intra macroblocks
    std::set<std::pair<int, int>> pairs, pairs_intra;
};
std::map<std::tuple<const void*, uint64_t, const void*, uint64_t, uint32_t>, H264Scan> g_h264_scans;

const H264Scan& scan_h264(const mmr_recon_h264_macroblock* mbs, const mmr_recon_h264_block* blocks, uint32_t n) {
    const auto key = std::make_tuple((const void*)mbs, model::content_key(mbs), (const void*)blocks, model::content_key(blocks), n);
    auto it = g_h264_scans.find(key);
    if (it != g_h264_scans.end()) return it->second;
    H264Scan r;
    for (uint32_t i = 0; i < n; ++i) {
        const bool intra = mbs[i].flags & MMR_H264_MB_INTRA;
        r.intra |= intra;
        // inter macroblocks read a level record only with cbp != 0; intra ones may hold one without
        // (Intra16x16 DC, I_PCM samples), so every intra macroblock's record index counts
        if (intra) r.max_coef_intra = std::max<int64_t>(r.max_coef_intra, mbs[i].coef_offset);
        else if (mbs[i].cbp != 0) r.max_coef = std::max<int64_t>(r.max_coef, mbs[i].coef_offset);
        for (int k = 0; k < 16; ++k) {
            const mmr_recon_h264_block& b = blocks[(size_t)i * 16 + k];
            (intra ? r.pairs_intra : r.pairs).insert({b.ref_idx[0] + 1, b.ref_idx[1] + 1});
        }
    }
    if (g_h264_scans.size() > 4096) g_h264_scans.clear();
    return g_h264_scans.emplace(key, r).first->second;
}

// DV placement (dv.hip: dv_mb_coordinates + the stores at the end of dv_kernel): the largest byte
// offset written in each plane, over every block of a frame, computed once per system.
struct DvExtent {
    int64_t end[3] = {0, 0, 0};
    bool negative = false;
};
void dv_mb_coordinates(uint32_t system, int seq, int slot, int mb, int* ox, int* oy) {
    const int OFF[5] = {2, 6, 8, 0, 4};
    const int SHUF3[5] = {18, 9, 27, 0, 36};
    const int LSTART[5] = {9, 4, 13, 0, 18};
    if (system == MMR_DV_SYSTEM_625_50) {
        const int serpent1 = (slot % 6 < 3) ? (slot % 6) : (5 - slot % 6);
        const int x = SHUF3[mb] + slot / 3;
        const int y = serpent1 + ((seq + OFF[mb]) % 12) * 3;
        *ox = x * 2;
        *oy = y * 2;
    } else {
        const int row = (seq + OFF[mb]) % 10;
        const int k = slot + ((mb == 1 || mb == 2) ? 3 : 0);
        const int x = LSTART[mb] + k / 6;
        const int serpent2 = (k % 12 < 6) ? (k % 12) : (11 - k % 12);
        int y = serpent2 + row * 6;
        if (x > 21) y = y * 2 - row * 6;
        *ox = x * 4;
        *oy = y;
    }
}
DvExtent dv_extent(uint32_t system) {
    static std::map<uint32_t, DvExtent> memo;
    auto it = memo.find(system);
    if (it != memo.end()) return it->second;
    DvExtent e;
    const uint32_t blocks = (system == MMR_DV_SYSTEM_625_50 ? 12u : 10u) * 27u * 5u * 6u;
    auto put = [&e](int plane, int64_t off, int bytes) {
        if (off < 0) e.negative = true;
        e.end[plane] = std::max(e.end[plane], off + bytes);
    };
    for (uint32_t blk = 0; blk < blocks; ++blk) {
        const uint32_t mb_index = blk / 6;
        const int b = (int)(blk % 6);
        const int seq = (int)(mb_index / 135), slot = (int)((mb_index % 135) / 5), mbs = (int)(mb_index % 5);
        int mbx, mby;
        dv_mb_coordinates(system, seq, slot, mbs, &mbx, &mby);
        const int X = mbx * 8, Y = mby * 8;
        for (int r = 0; r < 8; ++r) {
            if (system == MMR_DV_SYSTEM_625_50) {
                if (b < 4) put(0, (int64_t)(Y + (b >> 1) * 8 + r) * 720 + X + (b & 1) * 8, 8);
                else put(b == 4 ? 2 : 1, (int64_t)(Y / 2 + r) * 360 + X / 2, 8);
            } else if (mbx < 88) {
                if (b < 4) put(0, (int64_t)(Y + r) * 720 + X + b * 8, 8);
                else put(b == 4 ? 2 : 1, (int64_t)(Y + r) * 180 + X / 4, 8);
            } else {
                if (b < 4) {
                    put(0, (int64_t)(Y + (b >> 1) * 8 + r) * 720 + X + (b & 1) * 8, 8);
                } else {
                    const int64_t p = (int64_t)(Y + r) * 180 + X / 4;
                    put(b == 4 ? 2 : 1, p, 4);
                    put(b == 4 ? 2 : 1, p + 8 * 180, 4);
                }
            }
        }
    }
    memo.emplace(system, e);
    return e;
}

// find_job over ascending begin offsets (recon_internal.h): index of the last job whose begin <= key.
template <typename Get>
uint32_t find_job_host(uint32_t count, uint32_t key, Get get) {
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (get(mid) <= key) lo = mid; else hi = mid;
    }
    return lo;
}

// The reference pairs a list of blocks resolves to after h264_mb_kernel's clamp of invalid indices.
void h264_used_refs(const std::set<std::pair<int, int>>& pairs, const mmr::H264Job& job, bool used[2][MMR_H264_MAX_REFS]) {
    for (auto [a, b] : pairs) {
        int r0 = a - 1, r1 = b - 1;
        if (r0 >= (int)job.ref_count[0] || r1 >= (int)job.ref_count[1] || (r0 < 0 && r1 < 0)) {
            r0 = job.ref_count[0] ? std::min(std::max(r0, 0), (int)job.ref_count[0] - 1) : -1;
            r1 = (r0 < 0 && job.ref_count[1]) ? std::min(std::max(r1, 0), (int)job.ref_count[1] - 1) : -1;
            if (r0 < 0 && r1 < 0) r0 = 0;
        }
        if (r0 >= 0) used[0][r0] = true;
        if (r1 >= 0) used[1][r1] = true;
    }
}

void check_h264_planes(const mmr::H264Job& job, const char* kernel) {
    const size_t W = (size_t)job.width_mbs * 16, H = (size_t)job.height_mbs * 16;
    check(job.planes[0], W * H, kernel);
    check(job.planes[1], W * H / 4, kernel);
    check(job.planes[2], W * H / 4, kernel);
}

}  // namespace

namespace mmr {

hipError_t launch_fill_zero(uint32_t* d_words, uint32_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    model::enqueue(s, "fill_zero", [=] {
        if (check(d_words, (size_t)count * 4, "fill_zero words")) memset(d_words, 0, (size_t)count * 4);
    });
    return hipSuccess;
}

hipError_t dv_upload_tables(hipStream_t) { return hipSuccess; }

uint32_t h264_intra_band_rows() { return 8; }   // kIntraRows (h264_intra.hip)

bool h264_deblock_rows_supported(uint32_t rows) { return rows == 4 || rows == 8 || rows == 16; }

// mjpeg.hip: mjpeg_blocks_kernel. Block b = (tile - tile_begin) * BPG + g of the job is active when
// b < block_count; it reads levels + 64 b (128 B) and quant + 64 qsel[comp] (128 B) and writes
// rows y < plane_h, columns x0 .. min(x0 + 8, plane_w) of its component's plane (stride plane_w).
hipError_t launch_mjpeg(const MjpegJob* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t group_blocks,
                        hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    if (!group_blocks_ok(group_blocks)) return hipErrorInvalidValue;
    model::enqueue(s, "mjpeg", [=] {
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, MjpegJob> recs;
        if (!walk_tiles("mjpeg job / tile table", d_jobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        for (auto& [j, v] : offs) {
            const MjpegJob& job = recs[j];
            model::stats().jobs++;
            check_tiles("mjpeg", j, v, (job.block_count + group_blocks - 1) / group_blocks);
            if (!job.block_count) continue;
            check(job.levels, (size_t)job.block_count * 128, "mjpeg levels");
            bool used[4] = {false, false, false, false};
            uint32_t qmax = 0;
            for (uint32_t k = 0; k < job.blocks_per_mcu && k < 16; ++k) {
                const uint32_t c = job.blk_comp[k];
                if (c > 3) {
                    violation("mjpeg: block %u of the MCU names component %u", k, c);
                    continue;
                }
                used[c] = true;
                qmax = std::max<uint32_t>(qmax, job.qsel[c]);
            }
            check(job.quant, (size_t)(qmax + 1) * 64 * 2, "mjpeg quantisation tables");
            for (int c = 0; c < 4; ++c)
                if (used[c]) check(job.planes[c], (size_t)job.plane_w[c] * job.plane_h[c], "mjpeg plane");
        }
    });
    return hipSuccess;
}

// color.hip: yuv_to_rgb_kernel. Tile t belongs to job tile_job[t]; its four waves are wave numbers
// (t - tile_begin) * 4 .. + 3 of the job, wave w covers segments 4w .. 4w + 3 (set_rgb_layout:
// raster-order runs of 256 pixels, or 256-pixel pieces of rows). Row y reads Y row y, chroma row
// min(floor(y ch / height), ch - 1), chroma samples up to min(floor((width - 1) cw / width), cw - 1),
// and writes packed RGB / RGBA row y.
hipError_t launch_yuv_to_rgb(const RgbJob* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    model::enqueue(s, "yuv_to_rgb", [=] {
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, RgbJob> recs;
        if (!walk_tiles("yuv_to_rgb job / tile table", d_jobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        for (auto& [i, v] : offs) {
            const RgbJob& j = recs[i];
            model::stats().jobs++;
            RgbJob layout = j;
            set_rgb_layout(layout);
            if (j.waves_per_row != layout.waves_per_row || j.flat != layout.flat || j.segments != layout.segments)
                violation("yuv_to_rgb: job %u segment layout (%u per row, flat %u, %u segments) does not match width %u height %u", i,
                          j.waves_per_row, j.flat, j.segments, j.width, j.height);
            check_tiles("yuv_to_rgb", i, v, rgb_job_tiles(j));
            if (!j.width || !j.height || !j.cw || !j.ch) {
                violation("yuv_to_rgb: job %u has a zero dimension", i);
                continue;
            }
            if (j.cw > j.width || j.width > 65535 || j.height > 65535)
                violation("yuv_to_rgb: job %u sizes outside the kernel's range", i);
            if (j.inv_width != 1.0 / j.width || j.inv_height != 1.0 / j.height)
                violation("yuv_to_rgb: job %u reciprocals do not match its sizes", i);
            if (j.half_chroma != ((uint64_t)j.cw * 2 == j.width ? 1u : 0u) || j.limited != (j.range == MMR_RANGE_LIMITED ? 1u : 0u))
                violation("yuv_to_rgb: job %u chroma sharing / range flags do not match its sizes and range", i);
            const uint64_t ymax = j.height - 1;
            const uint64_t cy = std::min<uint64_t>(ymax * j.ch / j.height, j.ch - 1);
            const uint64_t cx = std::min<uint64_t>((uint64_t)(j.width - 1) * j.cw / j.width, j.cw - 1);
            check(j.planes[0], ymax * j.luma_stride + j.width, "yuv_to_rgb luma plane");
            check(j.planes[1], cy * j.chroma_stride + cx + 1, "yuv_to_rgb Cb plane");
            check(j.planes[2], cy * j.chroma_stride + cx + 1, "yuv_to_rgb Cr plane");
            if (j.rgb) check(j.rgb, (ymax + 1) * j.width * 3, "yuv_to_rgb RGB output");
            if (j.rgba) check(j.rgba, (ymax + 1) * j.width * 4, "yuv_to_rgb RGBA output");
        }
    });
    return hipSuccess;
}

// color.hip: linear_composite_kernel. Tile t belongs to job tile_job[t]; wave w of the job covers
// row w / waves_per_row, pixels (w % waves_per_row) * 64 .. + 63 clipped to the width. Each layer
// reads the pixels of the intersection of its clip, the target and its own extent.
hipError_t launch_linear_composite(const LinearCompositeDev* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles,
                                  const uint16_t* d_to_linear, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    model::enqueue(s, "linear_composite", [=] {
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, LinearCompositeDev> recs;
        if (!walk_tiles("linear_composite job / tile table", d_jobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        bool srgb = false;
        for (auto& [i, v] : offs) {
            const LinearCompositeDev& j = recs[i];
            model::stats().jobs++;
            if (j.waves_per_row != (j.width + 63) / 64) violation("linear_composite: job %u waves per row", i);
            check_tiles("linear_composite", i, v, linear_job_tiles(j.waves_per_row, j.height));
            if (!j.width || !j.height) continue;
            check(j.target, ((uint64_t)(j.height - 1) * j.stride + j.width) * 8, "linear_composite target");
            for (uint32_t l = 0; l < j.layer_count && l < MMR_RECON_MAX_LAYERS; ++l) {
                const mmr_recon_linear_layer& L = j.layers[l];
                const int64_t y0 = std::max<int64_t>({0, L.clip_top, L.y});
                const int64_t y1 = std::min<int64_t>({(int64_t)j.height, L.clip_bottom, (int64_t)L.y + L.height});
                const int64_t x0 = std::max<int64_t>({0, L.clip_left, L.x});
                const int64_t x1 = std::min<int64_t>({(int64_t)j.width, L.clip_right, (int64_t)L.x + L.width});
                if (y0 >= y1 || x0 >= x1) continue;
                const uint64_t bpp = L.format == MMR_LINEAR_SOURCE_SRGBA8 ? 4 : 8;
                srgb |= L.format == MMR_LINEAR_SOURCE_SRGBA8;
                check(L.pixels, ((uint64_t)(y1 - 1 - L.y) * L.stride + (uint64_t)(x1 - 1 - L.x) + 1) * bpp, "linear_composite layer");
            }
        }
        if (srgb) check(d_to_linear, 512, "linear_composite srgb_to_linear table");
    });
    return hipSuccess;
}

// color.hip: linear_convert_kernel, the same pixel walk; reads / writes one linear and one sRGBA8
// pixel per lane and the tables.
hipError_t launch_linear_convert(const LinearConvertDev* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles,
                                const uint8_t* d_tables, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    model::enqueue(s, "linear_convert", [=] {
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, LinearConvertDev> recs;
        if (!walk_tiles("linear_convert job / tile table", d_jobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        check(d_tables, kLinearTableBytes, "linear_convert tables");
        for (auto& [i, v] : offs) {
            const LinearConvertDev& j = recs[i];
            model::stats().jobs++;
            check_tiles("linear_convert", i, v, linear_job_tiles(j.waves_per_row, j.height));
            if (!j.width || !j.height) continue;
            check(j.linear, ((uint64_t)(j.height - 1) * j.linear_stride + j.width) * 8, "linear_convert linear surface");
            check(j.rgba, (uint64_t)(j.height - 1) * j.rgba_stride + (uint64_t)j.width * 4, "linear_convert sRGBA8 image");
        }
    });
    return hipSuccess;
}

// color.hip: composite_kernel. Tile t belongs to job tile_job[t]; wave w of the job covers pixels
// (w % waves_per_row) 256 .. +255 of row w / waves_per_row (rows below height, pixels below width):
// the target's rows and each layer's rows and columns that overlap them.
hipError_t launch_composite(const CompositeJobDev* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    model::enqueue(s, "composite", [=] {
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, CompositeJobDev> recs;
        if (!walk_tiles("composite job / tile table", d_jobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        for (auto& [i, v] : offs) {
            const CompositeJobDev& j = recs[i];
            model::stats().jobs++;
            if (j.waves_per_row != (j.width + 255) / 256) violation("composite: job %u waves per row", i);
            check_tiles("composite", i, v, composite_job_tiles(j.waves_per_row, j.height));
            if (!j.width || !j.height) continue;
            check(j.target, (uint64_t)(j.height - 1) * j.stride + (uint64_t)j.width * 4, "composite target");
            for (uint32_t l = 0; l < j.layer_count && l < MMR_RECON_MAX_LAYERS; ++l) {
                const mmr_recon_layer& L = j.layers[l];
                if (L.y >= j.height || L.height == 0 || L.width == 0 || L.x >= j.width) continue;
                const uint64_t rows = std::min<uint64_t>(L.height, j.height - L.y);
                const uint64_t cols = std::min<uint64_t>(L.width, j.width - L.x);
                check(L.rgba, (rows - 1) * L.stride + cols * 4, "composite layer");
            }
        }
    });
    return hipSuccess;
}

// color.hip: yuv_composite_kernel, composite_kernel's tiles over the composite jobs; job i also
// reads conversion record i and its planes (yuv_to_rgb's reads) and writes the target only.
hipError_t launch_yuv_composite(const RgbJob* d_rjobs, const CompositeJobDev* d_cjobs, const uint32_t* d_tile_job, uint32_t tile_base,
                                uint32_t tiles, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    model::enqueue(s, "yuv_composite", [=] {
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, CompositeJobDev> recs;
        if (!walk_tiles("yuv_composite job / tile table", d_cjobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        for (auto& [i, v] : offs) {
            const CompositeJobDev& j = recs[i];
            model::stats().jobs++;
            if (j.waves_per_row != (j.width + 255) / 256) violation("yuv_composite: job %u waves per row", i);
            check_tiles("yuv_composite", i, v, composite_job_tiles(j.waves_per_row, j.height));
            if (!check(d_rjobs + i, sizeof(RgbJob), "yuv_composite conversion record")) continue;
            const RgbJob& r = d_rjobs[i];
            if (!j.width || !j.height || !r.cw || !r.ch || r.width != j.width || r.height != j.height || j.stride != j.width * 4 ||
                r.rgba != j.target || r.rgb) {
                violation("yuv_composite: job %u conversion and composite do not describe one packed RGBA target", i);
                continue;
            }
            if (r.inv_width != 1.0 / r.width || r.inv_height != 1.0 / r.height || r.cw > r.width ||
                r.half_chroma != ((uint64_t)r.cw * 2 == r.width ? 1u : 0u) || r.limited != (r.range == MMR_RANGE_LIMITED ? 1u : 0u))
                violation("yuv_composite: job %u conversion record does not match its sizes and range", i);
            const uint64_t ymax = r.height - 1;
            const uint64_t cy = std::min<uint64_t>(ymax * r.ch / r.height, r.ch - 1);
            const uint64_t cx = std::min<uint64_t>((uint64_t)(r.width - 1) * r.cw / r.width, r.cw - 1);
            check(r.planes[0], ymax * r.luma_stride + r.width, "yuv_composite luma plane");
            check(r.planes[1], cy * r.chroma_stride + cx + 1, "yuv_composite Cb plane");
            check(r.planes[2], cy * r.chroma_stride + cx + 1, "yuv_composite Cr plane");
            check(j.target, (uint64_t)(j.height - 1) * j.stride + (uint64_t)j.width * 4, "yuv_composite target");
            for (uint32_t l = 0; l < j.layer_count && l < MMR_RECON_MAX_LAYERS; ++l) {
                const mmr_recon_layer& L = j.layers[l];
                if (L.y >= j.height || L.height == 0 || L.width == 0 || L.x >= j.width) continue;
                const uint64_t rows = std::min<uint64_t>(L.height, j.height - L.y);
                const uint64_t cols = std::min<uint64_t>(L.width, j.width - L.x);
                check(L.rgba, (rows - 1) * L.stride + cols * 4, "yuv_composite layer");
            }
        }
    });
    return hipSuccess;
}

// color.hip: yuv_blend_kernel (job blockIdx.y, pixel p < width * height).
hipError_t launch_yuv_blend(const mmr_recon_yuv_blend_job* d_jobs, uint32_t job_count, uint32_t, hipStream_t s) {
    if (job_count == 0) return hipSuccess;
    model::enqueue(s, "yuv_blend", [=] {
        if (!check(d_jobs, sizeof(mmr_recon_yuv_blend_job) * job_count, "yuv_blend job records")) return;
        for (uint32_t i = 0; i < job_count; ++i) {
            const mmr_recon_yuv_blend_job& j = d_jobs[i];
            model::stats().jobs++;
            if (!j.width || !j.height) continue;
            const uint64_t w = j.width, h = j.height;
            check(j.target[0], (j.y + h - 1) * (uint64_t)j.target_stride[0] + j.x + w, "yuv_blend target Y");
            for (int k = 1; k < 3; ++k)
                check(j.target[k], (j.y / 2 + h / 2 - 1) * (uint64_t)j.target_stride[k] + j.x / 2 + w / 2, "yuv_blend target chroma");
            check(j.overlay[0], w * h, "yuv_blend overlay Y");
            check(j.overlay[1], (w / 2) * (h / 2), "yuv_blend overlay U");
            check(j.overlay[2], (w / 2) * (h / 2), "yuv_blend overlay V");
            check(j.rgba, w * h * 4, "yuv_blend alpha");
            check(j.chroma_alpha, (w / 2) * (h / 2), "yuv_blend chroma alpha");
        }
    });
    return hipSuccess;
}

// color.hip: overlay_prepare_kernel (one thread per 2x2 cell).
hipError_t launch_overlay_prepare(const mmr_recon_overlay_prepare_job* d_jobs, uint32_t job_count, uint32_t, hipStream_t s) {
    if (job_count == 0) return hipSuccess;
    model::enqueue(s, "overlay_prepare", [=] {
        if (!check(d_jobs, sizeof(mmr_recon_overlay_prepare_job) * job_count, "overlay job records")) return;
        for (uint32_t i = 0; i < job_count; ++i) {
            const mmr_recon_overlay_prepare_job& j = d_jobs[i];
            model::stats().jobs++;
            const uint64_t px = (uint64_t)j.width * j.height, cells = (uint64_t)((j.width + 1) / 2) * ((j.height + 1) / 2);
            check(j.rgba, px * 4, "overlay RGBA");
            for (uint8_t* p : {j.y_limited, j.y_full})
                if (p) check(p, px, "overlay luma output");
            for (uint8_t* p : {j.u_limited, j.v_limited, j.u_full, j.v_full, j.chroma_alpha})
                if (p) check(p, cells, "overlay chroma output");
        }
    });
    return hipSuccess;
}

// color.hip: scale kernels; per plane the source window, the canvas and the f32 intermediate.
hipError_t launch_scale(const ScalePlaneDev* d_planes, uint32_t plane_count, uint32_t, uint32_t, uint32_t, hipStream_t s) {
    if (plane_count == 0) return hipSuccess;
    model::enqueue(s, "scale", [=] {
        if (!check(d_planes, sizeof(ScalePlaneDev) * plane_count, "scale plane records")) return;
        for (uint32_t i = 0; i < plane_count; ++i) {
            const ScalePlaneDev& d = d_planes[i];
            model::stats().jobs++;
            check(d.src, (uint64_t)(d.src_h - 1) * d.src_stride + d.src_w, "scale source");
            check(d.dst, (uint64_t)(d.dst_h - 1) * d.dst_stride + d.dst_w, "scale canvas");
            if (d.resize && !d.tiled) check(d.tmp, (uint64_t)d.src_w * d.copy_h * 4, "scale intermediate");
            if (d.resize) {
                check(d.vleft, (uint64_t)d.copy_h * 4, "scale vertical taps");
                check(d.hleft, (uint64_t)d.copy_w * 4, "scale horizontal taps");
                check(d.vw, (uint64_t)d.copy_h * d.vmax * 4, "scale vertical weights");
                check(d.hw, (uint64_t)d.copy_w * d.hmax * 4, "scale horizontal weights");
            }
        }
    });
    return hipSuccess;
}

// mpeg2.hip: mpeg2_order_kernel. Global row grow maps to the last job with row_begin <= grow and
// its row r = grow - row_begin: reads the row's 16-byte macroblock records (and bytes 8..11 of the
// field records of field-predicted ones) and writes 6 mb_w 16-byte entries at order + 6 mb_w r.
hipError_t launch_mpeg2_order(const Mpeg2Job* d_jobs, uint32_t job_count, uint32_t total_rows, uint32_t* d_status, hipStream_t s) {
    if (job_count == 0 || total_rows == 0) return hipSuccess;
    model::enqueue(s, "mpeg2_order", [=] {
        check(d_status, 4, "mpeg2_order status word");
        if (!check(d_jobs, sizeof(Mpeg2Job) * job_count, "mpeg2_order job records")) return;
        uint64_t rows = 0;
        for (uint32_t k = 0; k < job_count; ++k) {
            const Mpeg2Job& j = d_jobs[k];
            if (j.row_begin != rows) violation("mpeg2_order: job %u starts at row %u, the rows before it end at %llu", k, j.row_begin, (unsigned long long)rows);
            rows += j.mb_h;
            const uint64_t n = (uint64_t)j.mb_w * j.mb_h;
            if (!check(j.mbs, n * sizeof(mmr_recon_mpeg2_macroblock), "mpeg2_order macroblock records")) continue;
            const Mpeg2Scan sc = scan_mpeg2(j.mbs, (uint32_t)n);
            if (sc.field && j.field_motion) check(j.field_motion, n * sizeof(mmr_recon_mpeg2_field_motion), "mpeg2_order field records");
            check(j.order, n * 6 * 16, "mpeg2_order block order (scratch)");
        }
        if (rows != total_rows)
            violation("mpeg2_order: launched over %u rows, the jobs hold %llu (rows past the last job index past its records)", total_rows,
                      (unsigned long long)rows);
    });
    return hipSuccess;
}

// mpeg2.hip: mpeg2_kernel. Block pos = (tile - tile_begin) BPG + g of the picture's block order
// (active when pos < 6 mbs; inactive blocks read entry 0); it reads its plane record, the 256-byte
// matrices, levels + 64 index (coded) and writes rows of dst; fetch_ref reads rows clamped to the
// plane (field rows: sel + 2 r) and, for a direction that is not used, 12 bytes at the plane start.
hipError_t launch_mpeg2(const Mpeg2Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t group_blocks,
                        uint32_t* d_status, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    if (!group_blocks_ok(group_blocks)) return hipErrorInvalidValue;
    model::enqueue(s, "mpeg2", [=] {
        check(d_status, 4, "mpeg2 status word");
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, Mpeg2Job> recs;
        if (!walk_tiles("mpeg2 job / tile table", d_jobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        for (auto& [k, v] : offs) {
            const Mpeg2Job& j = recs[k];
            model::stats().jobs++;
            const uint64_t n = (uint64_t)j.mb_w * j.mb_h;
            check_tiles("mpeg2", k, v, (n * 6 + group_blocks - 1) / group_blocks);
            if (j.tile_count != (n * 6 + group_blocks - 1) / group_blocks) violation("mpeg2: job %u tile_count %u", k, j.tile_count);
            check(j.order, n * 6 * 16, "mpeg2 block order (scratch)");
            check(j.matrices, 256, "mpeg2 matrices");
            if (!check(j.mbs, n * sizeof(mmr_recon_mpeg2_macroblock), "mpeg2 macroblock records")) continue;
            const Mpeg2Scan sc = scan_mpeg2(j.mbs, (uint32_t)n);
            if (sc.max_level >= 0) check(j.levels, (size_t)(sc.max_level + 1) * 128, "mpeg2 levels");
            if (sc.field && j.field_motion) check(j.field_motion, n * sizeof(mmr_recon_mpeg2_field_motion), "mpeg2 field records");
            for (int p = 0; p < 3; ++p) {
                const Mpeg2Plane& pl = j.plane[p];
                const uint32_t need_w = p ? j.mb_w * 8 : j.mb_w * 16, need_h = p ? j.mb_h * 8 : j.mb_h * 16;
                if (pl.stride < need_w || pl.height < need_h)
                    violation("mpeg2: job %u plane %d is %ux%u, its macroblocks cover %ux%u", k, p, pl.stride, pl.height, need_w, need_h);
                const size_t bytes = (size_t)pl.stride * pl.height;
                check(pl.dst, bytes, "mpeg2 output plane");
                check(pl.fwd, 12, "mpeg2 forward plane (dummy row load)");
                check(pl.bwd, 12, "mpeg2 backward plane (dummy row load)");
                if (sc.fwd) check(pl.fwd, bytes, "mpeg2 forward reference plane");
                if (sc.bwd) check(pl.bwd, bytes, "mpeg2 backward reference plane");
            }
        }
    });
    return hipSuccess;
}

// dv.hip: dv_kernel. The blocks (tile - tile_begin) BPG + i, i < BPG, below block_count (taken by
// the tile's groups in mode-sorted order) read info[blk] and levels + 64 blk and store at their
// macroblock_coordinates position (dv_extent).
hipError_t launch_dv(const DvJob* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t group_blocks,
                     hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    if (!group_blocks_ok(group_blocks)) return hipErrorInvalidValue;
    model::enqueue(s, "dv", [=] {
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, DvJob> recs;
        if (!walk_tiles("dv job / tile table", d_jobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        for (auto& [k, v] : offs) {
            const DvJob& j = recs[k];
            model::stats().jobs++;
            check_tiles("dv", k, v, (j.block_count + group_blocks - 1) / group_blocks);
            const uint32_t expect = (j.system == MMR_DV_SYSTEM_625_50 ? 12u : 10u) * 27u * 5u * 6u;
            if (j.block_count != expect) violation("dv: job %u has %u blocks, the system has %u", k, j.block_count, expect);
            check(j.levels, (size_t)j.block_count * 128, "dv levels");
            check(j.info, j.block_count, "dv block info");
            const DvExtent e = dv_extent(j.system);
            if (e.negative) violation("dv: a placement of system %u lies before its plane", j.system);
            const size_t h = j.system == MMR_DV_SYSTEM_625_50 ? 576 : 480;
            const size_t cbytes = j.system == MMR_DV_SYSTEM_625_50 ? 360 * 288 : 180 * 480;
            if ((size_t)e.end[0] > 720 * h || (size_t)e.end[1] > cbytes || (size_t)e.end[2] > cbytes)
                violation("dv: placement of system %u reaches past the frame's planes", j.system);
            check(j.planes[0], (size_t)e.end[0], "dv luma plane");
            check(j.planes[1], (size_t)e.end[1], "dv Cb plane");
            check(j.planes[2], (size_t)e.end[2], "dv Cr plane");
        }
    });
    return hipSuccess;
}

// h264.hip: h264_mb_kernel. Macroblock (tile - tile_begin) 4 + wave (active when < mbs; inactive
// waves read record 0) reads its record, its 16 block records, levels + 384 coef_offset when
// cbp != 0, the weight tables, the scaling lists, the reference planes its blocks name (windows
// clamped into the plane) and writes its samples of the three planes.
hipError_t launch_h264_mc(const H264Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, uint32_t* d_status,
                          hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    model::enqueue(s, "h264_mc", [=] {
        check(d_status, 4, "h264_mc status word");
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, H264Job> recs;
        if (!walk_tiles("h264_mc job / tile table", d_jobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        for (auto& [k, v] : offs) {
            const H264Job& j = recs[k];
            model::stats().jobs++;
            const uint64_t n = (uint64_t)j.width_mbs * j.height_mbs;
            check_tiles("h264_mc", k, v, (n + 3) / 4);
            if (!check(j.mbs, n * sizeof(mmr_recon_h264_macroblock), "h264_mc macroblock records")) continue;
            if (!check(j.blocks, n * 16 * sizeof(mmr_recon_h264_block), "h264_mc block records")) continue;
            const H264Scan& sc = scan_h264(j.mbs, j.blocks, (uint32_t)n);
            // intra macroblocks of a picture with intra records leave before any other read
            const int64_t max_coef = j.intra ? sc.max_coef : std::max(sc.max_coef, sc.max_coef_intra);
            if (max_coef >= 0) {
                check(j.levels, (size_t)(max_coef + 1) * MMR_H264_LEVELS_PER_MB * 2, "h264_mc levels");
                check(j.scaling, 6 * 16 + 2 * 64, "h264_mc scaling lists");
            }
            if (j.weighted == MMR_H264_WEIGHTED_TABLE) {
                check(j.weights, sizeof(mmr_recon_h264_weight) * 2 * MMR_H264_MAX_REFS * 3, "h264_mc weights");
                check(j.bi_weights, sizeof(mmr_recon_h264_bi_weight) * MMR_H264_MAX_REFS * MMR_H264_MAX_REFS * 3, "h264_mc bi-weights");
            }
            check_h264_planes(j, "h264_mc output plane");
            bool used[2][MMR_H264_MAX_REFS] = {};
            h264_used_refs(sc.pairs, j, used);
            if (!j.intra) h264_used_refs(sc.pairs_intra, j, used);
            const size_t W = (size_t)j.width_mbs * 16, H = (size_t)j.height_mbs * 16;
            for (int list = 0; list < 2; ++list)
                for (int r = 0; r < (int)MMR_H264_MAX_REFS; ++r)
                    if (used[list][r]) {
                        check(j.ref_planes[list][r][0], W * H, "h264_mc luma reference");
                        check(j.ref_planes[list][r][1], W * H / 4, "h264_mc Cb reference");
                        check(j.ref_planes[list][r][2], W * H / 4, "h264_mc Cr reference");
                    }
        }
    });
    return hipSuccess;
}

// deblock.hip: h264_bs_kernel. Macroblock a = (tile - tile_begin) 4 + group of a deblocked picture
// reads its record and its left / top neighbour's, their block records, and writes bs + 16 a.
hipError_t launch_h264_bs(const H264Job* d_jobs, const uint32_t* d_tile_job, uint32_t tile_base, uint32_t tiles, hipStream_t s) {
    if (tiles == 0) return hipSuccess;
    model::enqueue(s, "h264_bs", [=] {
        std::map<uint32_t, std::vector<uint32_t>> offs;
        std::map<uint32_t, H264Job> recs;
        if (!walk_tiles("h264_bs job / tile table", d_jobs, d_tile_job, tile_base, tiles, offs, recs)) return;
        for (auto& [k, v] : offs) {
            const H264Job& j = recs[k];
            model::stats().jobs++;
            const uint64_t n = (uint64_t)j.width_mbs * j.height_mbs;
            check_tiles("h264_bs", k, v, (n + 3) / 4);
            if (!j.deblock) continue;
            check(j.mbs, n * sizeof(mmr_recon_h264_macroblock), "h264_bs macroblock records");
            check(j.blocks, n * 16 * sizeof(mmr_recon_h264_block), "h264_bs block records");
            check(j.bs, n * 16, "h264_bs strengths (scratch)");
        }
    });
    return hipSuccess;
}

// deblock.hip: h264_deblock_kernel. Band ticket t < total_bands maps to the last job with
// band_begin <= t; wave w filters row (t - band_begin) ROWS + w when the job is deblocked and the
// row exists, reading / writing the planes around its macroblocks, reading bs and the records;
// counters[0] is the ticket, counters[1 + t] band t's progress.
hipError_t launch_h264_deblock(const H264Job* d_jobs, uint32_t job_count, uint32_t total_bands, uint32_t* d_counters, uint32_t groups,
                               uint32_t band_rows, uint32_t publish, uint32_t* d_status, hipStream_t s) {
    if (job_count == 0 || total_bands == 0) return hipSuccess;
    if (!h264_deblock_rows_supported(band_rows)) return hipErrorInvalidValue;
    (void)groups;
    (void)publish;
    model::enqueue(s, "h264_deblock", [=] {
        check(d_status, 4, "h264_deblock status word");
        check(d_counters, ((size_t)total_bands + 1) * 4, "h264_deblock counters");
        if (!check(d_jobs, sizeof(H264Job) * job_count, "h264_deblock job records")) return;
        std::vector<std::vector<uint32_t>> rows(job_count);
        for (uint32_t t = 0; t < total_bands; ++t) {
            const uint32_t k = find_job_host(job_count, t, [&](uint32_t i) { return d_jobs[i].band_begin; });
            const H264Job& j = d_jobs[k];
            const uint32_t band = t - j.band_begin;
            for (uint32_t w = 0; w < band_rows; ++w) {
                const uint64_t y = (uint64_t)band * band_rows + w;
                if (j.deblock && y < j.height_mbs) rows[k].push_back((uint32_t)y);
            }
        }
        for (uint32_t k = 0; k < job_count; ++k) {
            const H264Job& j = d_jobs[k];
            if (!j.deblock) {
                if (!rows[k].empty()) violation("h264_deblock: picture %u is not deblocked but gets rows", k);
                continue;
            }
            model::stats().jobs++;
            std::sort(rows[k].begin(), rows[k].end());
            bool ok = rows[k].size() == j.height_mbs;
            for (size_t i = 0; ok && i < rows[k].size(); ++i) ok = rows[k][i] == i;
            if (!ok) violation("h264_deblock: picture %u gets %zu rows, it has %u", k, rows[k].size(), j.height_mbs);
            const uint64_t n = (uint64_t)j.width_mbs * j.height_mbs;
            check(j.mbs, n * sizeof(mmr_recon_h264_macroblock), "h264_deblock macroblock records");
            check(j.bs, n * 16, "h264_deblock strengths (scratch)");
            check_h264_planes(j, "h264_deblock plane");
        }
    });
    return hipSuccess;
}

// h264_intra.hip: h264_intra_kernel. Band ticket t maps to the last job with intra_band_begin <= t;
// wave w reconstructs the intra macroblocks of row band * 8 + w: records, intra records, levels,
// the planes (the macroblock and its neighbours).
hipError_t launch_h264_intra(const H264Job* d_jobs, uint32_t job_count, uint32_t total_bands, uint32_t* d_counters, uint32_t,
                             uint32_t, uint32_t* d_status, hipStream_t s) {
    if (job_count == 0 || total_bands == 0) return hipSuccess;
    model::enqueue(s, "h264_intra", [=] {
        check(d_status, 4, "h264_intra status word");
        check(d_counters, ((size_t)total_bands + 1) * 4, "h264_intra counters");
        if (!check(d_jobs, sizeof(H264Job) * job_count, "h264_intra job records")) return;
        std::vector<uint32_t> rows(job_count, 0);
        for (uint32_t t = 0; t < total_bands; ++t) {
            const uint32_t k = find_job_host(job_count, t, [&](uint32_t i) { return d_jobs[i].intra_band_begin; });
            const H264Job& j = d_jobs[k];
            const uint32_t band = t - j.intra_band_begin;
            for (uint32_t w = 0; w < 8; ++w)
                if (j.intra && band * 8 + w < j.height_mbs) rows[k]++;
        }
        for (uint32_t k = 0; k < job_count; ++k) {
            const H264Job& j = d_jobs[k];
            if (!j.intra) continue;
            model::stats().jobs++;
            if (rows[k] != j.height_mbs) violation("h264_intra: picture %u gets %u rows, it has %u", k, rows[k], j.height_mbs);
            const uint64_t n = (uint64_t)j.width_mbs * j.height_mbs;
            check(j.mbs, n * sizeof(mmr_recon_h264_macroblock), "h264_intra macroblock records");
            check(j.intra, n * sizeof(mmr_recon_h264_intra), "h264_intra intra records");
            if (!check(j.blocks, n * 16 * sizeof(mmr_recon_h264_block), "h264_intra block records")) continue;
            const H264Scan& sc = scan_h264(j.mbs, j.blocks, (uint32_t)n);
            if (sc.max_coef_intra >= 0) check(j.levels, (size_t)(sc.max_coef_intra + 1) * MMR_H264_LEVELS_PER_MB * 2, "h264_intra levels");
            check_h264_planes(j, "h264_intra plane");
        }
    });
    return hipSuccess;
}

}  // namespace mmr
