// Audit of the blend's sample masks (fh_mask_prep / fh_mask_row_w2 /
// fh_mask_row_bits on the ellipse record fh_ellipse_record that the
// preprocess stores; every sample for records without a usable ellipse): the
// blend drops a listed Gaussian whose mask holds no sample of the tile (the
// tile cull), and its pixels evaluate only the staged Gaussians whose masks
// hold their bit (the first pass). Over every (tile, instance) pair of a
// frame and every sample, with the blend's own float32 operations (cutoff,
// exp, alpha >= 1/255), a sample a Gaussian reaches must be in its mask: then
// a dropped Gaussian reaches no sample and a pixel skips no Gaussian that
// reaches it. The kernel takes the row half-widths' square roots with
// v_sqrt_f32; the audit takes them lowered by a relative 2^-20, far beyond
// that instruction's error, so its masks are subsets of the kernel's. Input is a binary dump of one oracle frame written
// by tests/test_cull.py; prints pair counts and the number of violations of
// each (both must be 0), and the smallest slack, in pixels, between a reached
// sample and its row interval's edge. The same check covers the masks of the
// 16x8 and 16x16 work tiles (work_tile_mask) on FOVEA and BLENDED tiles.
//
// With --waves it prints instead the blend's wave-level loop counts for the
// frame, as the kernel runs it (batches of 64 staged Gaussians per 8x8 wave,
// four 16-lane QUARTER groups of 16 each): staging rounds, first-pass steps
// (the wave's largest kept count per round) and second-pass steps (the
// largest number of Gaussians any lane of the wave evaluates per round),
// beside what larger batches and a 16x16 work tile sharing one staging for
// the four sub-tiles of a FOVEA or BLENDED tile would give.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <string>
#include <algorithm>
#include "../foveatedhtgscudabackend_amd/csrc/fhtgs_cull.h"

static float dot3(const float* a, const float* b) { return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0])); }
static void cross3(const float* a, const float* b, float* o) {
    o[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    o[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    o[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}

// alpha >= 1/255 at pixel (px, py): inference.cuh:202-222 as the blend computes it
static bool reaches(const float* V1, const float* V2, const float* V4, float opacity, float pxf, float pyf) {
    float Px[4], Py[4];
    for (int c = 0; c < 4; ++c) {
        Px[c] = fmaf(-V4[c], pxf, V1[c]);
        Py[c] = fmaf(-V4[c], pyf, V2[c]);
    }
    float m[3], d[3];
    for (int c = 0; c < 3; ++c) m[c] = fmaf(Px[3], Py[c], -(Px[c] * Py[3]));
    cross3(Px, Py, d);
    const float num = dot3(m, m), den = dot3(d, d);
    if (num > 11.0825270903f * den) return false;
    const float G = fh_expf((-0.5f * num) * (1.0f / den));
    return fh_fmin(opacity * G, 1.0f) >= 1.0f / 255.0f;
}

// The staged Gaussian's sample mask as the blend computes it for the samples
// (bx + c step, by + r step), c, r < side, with each row's square root
// lowered by a relative 2^-20 (see above). `all`: the record has no usable
// ellipse (every sample). Per row, the interval's edges in columns go to lo
// and hi (NaN for an empty row), for the slack report.
static uint64_t sample_mask(const float rec[5], int bx, int by, int step, int side, bool* all, float lo[8],
                            float hi[8], float uc_out[8]) {
    const float inv_step = 1.0f / (float)step;
    const FhMaskPrep p = fh_mask_prep(rec, 1.0f / rec[3], (float)bx, (float)by, inv_step);
    *all = p.all;
    if (p.all) return side == 8 ? ~0ull : 0xffffull;
    uint64_t m = 0;
    for (int r = 0; r < side; ++r) {
        const float dy = fmaf((float)r, (float)step, p.dy0);
        const float w2 = fh_mask_row_w2(p, dy);
        const float w = w2 < 0.0f ? -1.0f : sqrtf(w2) * (1.0f - 0x1p-20f) + kMaskMarginPx;
        m |= (uint64_t)fh_mask_row_bits(p, dy, w, side) << (r * side);
        const float uc = fmaf(p.s, dy, p.u0);
        uc_out[r] = uc;
        lo[r] = w < 0.0f ? NAN : uc - w;
        hi[r] = w < 0.0f ? NAN : uc + w;
    }
    return m;
}

// The mask of sub-tile (sx, sy) of a 16-column work tile whose first pixel
// is (wx, wy), with `rows` pixel rows (8: 16x8, 16: 16x16), as the wide blend
// computes it: one 16-column row interval per pixel row of the work tile
// (fh_mask_row_bits with 16 columns), its low 8 columns the left sub-tile's,
// its high 8 the right one's; bit r * 8 + c for the sub-tile's sample (c, r).
// Square roots lowered as in sample_mask.
static uint64_t work_tile_mask(const float rec[5], int wx, int wy, int rows, int sx, int sy) {
    const FhMaskPrep p = fh_mask_prep(rec, 1.0f / rec[3], (float)wx, (float)wy, 1.0f);
    if (p.all) return ~0ull;
    uint64_t m = 0;
    for (int r = 0; r < rows; ++r) {
        if (r / 8 != sy) continue;
        const float dy = (float)r + p.dy0;
        const float w2 = fh_mask_row_w2(p, dy);
        const float w = w2 < 0.0f ? -1.0f : sqrtf(w2) * (1.0f - 0x1p-20f) + kMaskMarginPx;
        const uint32_t bits = fh_mask_row_bits(p, dy, w, 16);
        m |= (uint64_t)((bits >> (8 * sx)) & 0xffu) << ((r % 8) * 8);
    }
    return m;
}

template <class T>
static bool rd(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

struct Frame {
    int64_t gw = 0, W = 0, H = 0;
    std::vector<uint32_t> part, tmp, ranges, vals;
    std::vector<float> v1, v2, v4, rgba;
};

// The sample positions of partitioned tile tg (8x8, or 4x4 at stride 4 for
// QUARTER), -1 outside the image, and their rectangle.
static std::vector<std::pair<int, int>> samples(const Frame& F, int type, uint32_t tg, float r[4]) {
    const int step = type == 3 ? 4 : (type == 1 ? 2 : 1), side = type == 3 ? 4 : 8;
    const uint32_t lt = tg / 4, st = tg % 4;
    const int bx = (int)(lt % F.gw) * 16 + (int)(st % 2) * 8, by = (int)(lt / F.gw) * 16 + (int)(st / 2) * 8;
    r[0] = (float)bx; r[1] = (float)(bx + (side - 1) * step);
    r[2] = (float)by; r[3] = (float)(by + (side - 1) * step);
    std::vector<std::pair<int, int>> px;
    for (int q = 0; q < side * side; ++q) {
        const int x = bx + (q % side) * step, y = by + (q / side) * step;
        px.push_back(x < F.W && y < F.H ? std::make_pair(x, y) : std::make_pair(-1, -1));
    }
    return px;
}

// One lane group walking list [b, e) in batches of bs, culled against rect
// `cull`, first-pass tested at `px` about the rect's centre: per batch the
// kept count and each sample's count of passing Gaussians.
static void walk(const Frame& F, const std::vector<std::pair<int, int>>& px, const float srect[4],
                 const float cull[4], uint32_t b, uint32_t e, uint32_t bs, std::vector<int>& kept,
                 std::vector<std::vector<int>>& lane) {
    for (uint32_t c = b; c < e; c = (e - c > bs) ? c + bs : e) {
        int k = 0;
        std::vector<int> cnt(px.size(), 0);
        const uint32_t stop = (e - c > bs) ? c + bs : e;
        for (uint32_t j = c; j < stop; ++j) {
            const uint32_t p = F.vals[j];
            const float *a1 = &F.v1[4 * (size_t)p], *a2 = &F.v2[4 * (size_t)p], *a4 = &F.v4[4 * (size_t)p];
            const float o = F.rgba[4 * (size_t)p + 3];
            float rec[5];
            fh_ellipse_record(a1, a2, a4, 2.0f * fh_logf(o * 255.0f), rec);
            // the mask over this lane group's samples (rectangle srect); the
            // tile cull keeps a Gaussian whose mask over the cull rectangle's
            // samples holds one (the same samples unless a 16x16 work tile
            // is estimated)
            const int side = px.size() == 16 ? 4 : 8;
            const int step = (int)(srect[1] - srect[0]) / (side - 1);
            bool all = false;
            float lo[8], hi[8], uc[8];
            const uint64_t m = sample_mask(rec, (int)srect[0], (int)srect[2], step, side, &all, lo, hi, uc);
            bool kept = m != 0;
            if (cull != srect) {
                kept = false;
                for (int sy = 0; sy < 2 && !kept; ++sy)
                    for (int sx = 0; sx < 2 && !kept; ++sx)
                        kept = sample_mask(rec, (int)cull[0] + 8 * sx, (int)cull[2] + 8 * sy, 1, 8, &all, lo, hi,
                                           uc) != 0;
            }
            if (!kept) continue;
            ++k;
            for (size_t q = 0; q < px.size(); ++q)
                if (px[q].first >= 0) cnt[q] += (int)((m >> q) & 1ull);
        }
        kept.push_back(k);
        lane.push_back(cnt);
    }
}

static int max_of(const std::vector<int>& v) { return v.empty() ? 0 : *std::max_element(v.begin(), v.end()); }

// second-pass steps when g consecutive batches are evaluated as one
static double merged(const std::vector<std::vector<int>>& lane, size_t g) {
    double s = 0;
    for (size_t i = 0; i < lane.size(); i += g) {
        std::vector<int> sum(lane.empty() ? 0 : lane[0].size(), 0);
        for (size_t j = i; j < lane.size() && j < i + g; ++j)
            for (size_t q = 0; q < sum.size(); ++q) sum[q] += lane[j][q];
        s += max_of(sum);
    }
    return s;
}

static void wave_report(const Frame& F) {
    const char* names[4] = {"FOVEA", "PERIPHERY", "BLENDED", "QUARTER"};
    double tot[3] = {0, 0, 0};
    for (int type = 0; type < 4; ++type) {
        double rounds = 0, p1 = 0, p2 = 0, alt[3] = {0, 0, 0}, wt_rounds = 0, wt_p1 = 0, units = 0;
        const uint32_t g0 = F.part[2 * type], g1 = F.part[2 * type + 1];
        if (type == 3) {
            // four tiles per wave, 16 lanes and 16 staged Gaussians each; the
            // wave runs as many rounds and steps as its largest group needs
            for (uint32_t g = g0; g < g1; g += 4) {
                ++units;
                for (int v = 0; v < 3; ++v) {
                    const uint32_t bs = v == 0 ? 16u : (v == 1 ? 32u : 0xffffffffu);
                    std::vector<std::vector<int>> kept(4);
                    std::vector<std::vector<std::vector<int>>> lane(4);
                    size_t nb = 0;
                    for (uint32_t t = 0; t < 4 && g + t < g1; ++t) {
                        float r[4];
                        const uint32_t tg = F.tmp[g + t];
                        const auto px = samples(F, type, tg, r);
                        walk(F, px, r, r, F.ranges[2 * (size_t)tg], F.ranges[2 * (size_t)tg + 1], bs, kept[t], lane[t]);
                        nb = std::max(nb, kept[t].size());
                    }
                    for (size_t i = 0; i < nb; ++i) {
                        int km = 0, pm = 0;
                        for (int t = 0; t < 4; ++t)
                            if (i < kept[t].size()) { km = std::max(km, kept[t][i]); pm = std::max(pm, max_of(lane[t][i])); }
                        if (v == 0) { rounds += 1; p1 += km; p2 += pm; }
                        alt[v] += pm;
                    }
                }
            }
            std::printf("waves %-9s units %.0f rounds %.0f pass1 %.0f pass2 %.0f | pass2 with 32 per group %.0f, "
                        "whole lists %.0f\n", names[type], units, rounds, p1, p2, alt[1], alt[2]);
        } else {
            for (uint32_t g = g0; g < g1; ++g) {
                ++units;
                float r[4];
                const uint32_t tg = F.tmp[g];
                const auto px = samples(F, type, tg, r);
                std::vector<int> kept;
                std::vector<std::vector<int>> lane;
                walk(F, px, r, r, F.ranges[2 * (size_t)tg], F.ranges[2 * (size_t)tg + 1], 64, kept, lane);
                rounds += kept.size();
                for (size_t i = 0; i < kept.size(); ++i) { p1 += kept[i]; p2 += max_of(lane[i]); }
                alt[0] += merged(lane, 2);
                alt[1] += merged(lane, 4);
                alt[2] += merged(lane, lane.size() + 1);
                if (type != 1 && tg % 4 == 0) {
                    // 16x16 work tile: the sub-tiles' lists are the same, staged
                    // once, culled against the whole tile, each wave's first
                    // pass over every kept Gaussian
                    const uint32_t lt = tg / 4;
                    const float big[4] = {(float)((lt % F.gw) * 16), (float)((lt % F.gw) * 16 + 15),
                                          (float)((lt / F.gw) * 16), (float)((lt / F.gw) * 16 + 15)};
                    for (uint32_t st = 0; st < 4; ++st) {
                        float rs[4];
                        const auto pxs = samples(F, type, lt * 4 + st, rs);
                        std::vector<int> k2;
                        std::vector<std::vector<int>> l2;
                        walk(F, pxs, rs, big, F.ranges[2 * (size_t)tg], F.ranges[2 * (size_t)tg + 1], 64, k2, l2);
                        if (st == 0) wt_rounds += k2.size();
                        for (int k : k2) wt_p1 += k;
                    }
                }
            }
            std::printf("waves %-9s units %.0f rounds %.0f pass1 %.0f pass2 %.0f | pass2 with batches of 128 %.0f, "
                        "256 %.0f, whole lists %.0f", names[type], units, rounds, p1, p2, alt[0], alt[1], alt[2]);
            if (type != 1) std::printf(" | 16x16 work tile: staging rounds %.0f, pass1 %.0f", wt_rounds, wt_p1);
            std::printf("\n");
        }
        tot[0] += rounds; tot[1] += p1; tot[2] += p2;
    }
    std::printf("waves total rounds %.0f pass1 %.0f pass2 %.0f\n", tot[0], tot[1], tot[2]);
}

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: cull_audit dump.bin [--waves]\n"); return 2; }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) { std::perror("open"); return 2; }
    int64_t h[7];
    if (fread(h, 8, 7, f) != 7) return 2;
    const int64_t N = h[0], I = h[1], n_small = h[2], n_active = h[3], gw = h[4], W = h[5], H = h[6];
    std::vector<uint32_t> part, tmp, ranges, vals;
    std::vector<float> v1, v2, v4, rgba;
    if (!rd(f, part, 8) || !rd(f, tmp, n_active) || !rd(f, ranges, 2 * n_small) || !rd(f, vals, I) ||
        !rd(f, v1, 4 * N) || !rd(f, v2, 4 * N) || !rd(f, v4, 4 * N) || !rd(f, rgba, 4 * N)) {
        std::fprintf(stderr, "short read\n");
        return 2;
    }
    std::fclose(f);
    if (argc > 2 && std::string(argv[2]) == "--waves") {
        Frame F;
        F.gw = gw; F.W = W; F.H = H;
        F.part = part; F.tmp = tmp; F.ranges = ranges; F.vals = vals;
        F.v1 = v1; F.v2 = v2; F.v4 = v4; F.rgba = rgba;
        wave_report(F);
        return 0;
    }
    int64_t pairs = 0, kept = 0, reach = 0, bad = 0;
    int64_t px_reach = 0, px_pre = 0, px_bad = 0, px_evals = 0, no_conic = 0;
    // FOVEA / BLENDED samples a Gaussian reaches, and those missing from its
    // 16x8 or 16x16 work-tile mask
    int64_t wt_reach = 0, wt_bad = 0, wt_bits8 = 0, wt_bits16 = 0;
    double min_slack = 1e30;
    for (int type = 0; type < 4; ++type) {
        // periphery tiles sample their large tile at stride 2 (8x8 samples),
        // quarter tiles at stride 4 (4x4 samples)
        const int step = type == 3 ? 4 : (type == 1 ? 2 : 1);
        const int side = type == 3 ? 4 : 8;
        for (uint32_t g = part[2 * type]; g < part[2 * type + 1]; ++g) {
            const uint32_t tg = tmp[g], lt = tg / 4, st = tg % 4;
            const int bx = (int)(lt % gw) * 16 + (int)(st % 2) * 8;
            const int by = (int)(lt / gw) * 16 + (int)(st / 2) * 8;
            for (uint32_t j = ranges[2 * (size_t)tg]; j < ranges[2 * (size_t)tg + 1]; ++j) {
                const uint32_t p = vals[j];
                const float *a1 = &v1[4 * (size_t)p], *a2 = &v2[4 * (size_t)p], *a4 = &v4[4 * (size_t)p];
                const float o = rgba[4 * (size_t)p + 3];
                float rec[5];
                fh_ellipse_record(a1, a2, a4, 2.0f * fh_logf(o * 255.0f), rec);
                bool all = false;
                float lo[8], hi[8], uc[8];
                const uint64_t m = sample_mask(rec, bx, by, step, side, &all, lo, hi, uc);
                const bool keep = m != 0;
                const bool wide = type == 0 || type == 2;
                uint64_t mw8 = 0, mw16 = 0;
                if (wide) {
                    const int wx = (int)(lt % gw) * 16, sx = (int)(st % 2), sy = (int)(st / 2);
                    mw8 = work_tile_mask(rec, wx, (int)(lt / gw) * 16 + 8 * sy, 8, sx, 0);
                    mw16 = work_tile_mask(rec, wx, (int)(lt / gw) * 16, 16, sx, sy);
                    wt_bits8 += __builtin_popcountll(mw8);
                    wt_bits16 += __builtin_popcountll(mw16);
                }
                bool any = false;
                for (int q = 0; q < side * side && !any; ++q) {
                    const int px = bx + (q % side) * step, py = by + (q / side) * step;
                    if (px < W && py < H) any = reaches(a1, a2, a4, o, (float)px, (float)py);
                }
                ++pairs;
                kept += keep;
                reach += any;
                if (keep) {
                    no_conic += all;
                    for (int q = 0; q < side * side; ++q) {
                        const int px = bx + (q % side) * step, py = by + (q / side) * step;
                        if (!(px < W && py < H)) continue;
                        const bool r = reaches(a1, a2, a4, o, (float)px, (float)py);
                        const bool pt = (m >> q) & 1ull;
                        ++px_evals;
                        px_reach += r;
                        px_pre += pt;
                        if (r && !all) {
                            // slack in pixels between the sample's column and its row interval's edges
                            const int c = q % side, rr = q / side;
                            const double sl = std::min((double)c - (double)lo[rr], (double)hi[rr] - (double)c) * step;
                            if (sl < min_slack) min_slack = sl;
                        }
                        if (wide && r) {
                            ++wt_reach;
                            if (!((mw8 >> q) & 1ull) || !((mw16 >> q) & 1ull)) {
                                if (wt_bad < 5)
                                    std::printf("work-tile violation: tile %u prim %u px %d py %d\n", tg, p, px, py);
                                ++wt_bad;
                            }
                        }
                        if (r && !pt) {
                            if (px_bad < 5)
                                std::printf("pixel violation: tile %u prim %u px %d py %d rec=(%a %a %a %a %a)\n", tg,
                                            p, px, py, rec[0], rec[1], rec[2], rec[3], rec[4]);
                            ++px_bad;
                        }
                    }
                }
                if (any && !keep) {
                    if (bad < 5)
                        std::printf("violation: tile %u prim %u a4=(%a %a %a %a) o=%a\n", tg, p, a4[0], a4[1],
                                    a4[2], a4[3], o);
                    ++bad;
                }
            }
        }
    }
    std::printf("pairs %lld kept %lld reaching %lld violations %lld\n", (long long)pairs, (long long)kept,
                (long long)reach, (long long)bad);
    std::printf("pixels %lld reached %lld pretest %lld pixel_violations %lld no_conic %lld min_slack %.6g\n",
                (long long)px_evals, (long long)px_reach, (long long)px_pre, (long long)px_bad, (long long)no_conic,
                min_slack);
    std::printf("work_tiles reached %lld violations %lld mask_bits_16x8 %lld mask_bits_16x16 %lld\n",
                (long long)wt_reach, (long long)wt_bad, (long long)wt_bits8, (long long)wt_bits16);
    return bad != 0 || px_bad != 0 || wt_bad != 0;
}
