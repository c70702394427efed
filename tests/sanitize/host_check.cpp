// host_check.cpp -- the engine's host-only code (tap-table builders, Java float text, kernel
// constants, rolling/load plan) run under AddressSanitizer and UndefinedBehaviorSanitizer,
// with libstdc++ bounds assertions, over wide parameter grids.  Besides running clean it
// checks properties the GPU kernels depend on:
//   * Float.toString text reads back to the same float, and the five BIC tSO texts of
//     LiveSRRF_CL.java:318 at fwhm 1.0-3.0;
//   * the sweep's union tap tables of several fwhm values reduce, per fwhm, to that fwhm's
//     own tables bit for bit (taps in order, weights, unit vectors, weight sums, bicubic);
//   * every BIC and 3-D tap offset lies inside the LDS tile the kernels allocate;
//   * 3-D tables at axial factor 3*Mz hold, for plane 3z+1, the taps of plane z at Mz;
//   * the rolling plan's loads cover each SR window's frames exactly once.
// Exit status 0 and the line "host_check: ok" mean every check passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/esrrf.h"
#include "../../nanoj-esrrf_amd/csrc/geometry.h"
#include "../../nanoj-esrrf_amd/csrc/java_text.h"

using namespace esrrf;

static int g_fail = 0;
#define CHECK(cond, ...)                                    \
    do {                                                    \
        if (!(cond)) {                                      \
            if (g_fail < 20) {                              \
                fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
                fprintf(stderr, __VA_ARGS__);               \
                fprintf(stderr, "\n");                      \
            }                                               \
            g_fail++;                                       \
        }                                                   \
    } while (0)

static bool same_bits(float a, float b) { return std::memcmp(&a, &b, sizeof a) == 0; }

static void check_java_text()
{
    struct K { float x; const char *s; };
    const K known[] = {{1.0f, "1.0"},       {0.001f, "0.001"},     {1.0e-4f, "1.0E-4"},
                       {1.0e7f, "1.0E7"},   {9999999.0f, "9999999.0"}, {100.0f, "100.0"},
                       {0.1f, "0.1"},       {-2.5f, "-2.5"},       {1.4e-45f, "1.4E-45"},
                       {3.4028235e38f, "3.4028235E38"}, {0.0f, "0.0"}, {123456.7f, "123456.7"}};
    for (const K &k : known) {
        const std::string t = java_float_to_string(k.x);
        CHECK(t == k.s, "Float.toString(%.9g) = %s, expected %s", (double)k.x, t.c_str(), k.s);
    }
    std::mt19937 rng(12345);
    int n = 0;
    while (n < 60000) {
        uint32_t bits = rng();
        float x;
        std::memcpy(&x, &bits, sizeof x);
        if (!std::isfinite(x)) continue;
        n++;
        const std::string t = java_float_to_string(x);
        const float back = opencl_literal_to_float(t);
        CHECK(same_bits(back, x) || (x == 0.0f && back == 0.0f),
              "round trip of %.9g via %s gives %.9g", (double)x, t.c_str(), (double)back);
    }
    const float fw[] = {1.0f, 1.5f, 2.0f, 2.5f, 3.0f};
    const char *tso[] = {"0.849617661", "1.27442651", "1.69923531", "2.12404421", "2.5488531"};
    for (int i = 0; i < 5; i++) {
        char buf[64];
        CHECK(esrrf_bic_tso_text(fw[i], buf, sizeof buf) == ESRRF_OK, "tso text");
        CHECK(std::strcmp(buf, tso[i]) == 0, "tSO text at fwhm %g: %s, expected %s",
              (double)fw[i], buf, tso[i]);
        float s, r, tss, t;
        esrrf_bic_constants(fw[i], &s, &r, &tss, &t);
        CHECK(same_bits(t, (float)std::strtod(tso[i], nullptr)), "tSO value at fwhm %g",
              (double)fw[i]);
    }
}

static void bic_consts(float fwhm, float &radius, float &tSS, float &tSO)
{
    float s;
    esrrf_bic_constants(fwhm, &s, &radius, &tSS, &tSO);
}

static void check_bic_tables()
{
    const int sizes[][2] = {{4, 4}, {5, 7}, {13, 9}, {33, 17}, {64, 40}};
    const float fwhms[] = {0.5f, 1.0f, 1.5f, 2.0f, 2.5f, 3.0f, 4.0f};
    long tables = 0;
    for (const auto &sz : sizes) {
        for (int M = 1; M <= 8; M++) {
            for (float fwhm : fwhms) {
                float r, tss, tso;
                bic_consts(fwhm, r, tss, tso);
                BicGeometry g;
                if (!build_bic_geometry(sz[0], sz[1], M, r, tss, tso, g)) continue;
                tables++;
                const long tile = 4L * g.PP * g.PH;
                for (const Phase &P : g.phases) {
                    CHECK(P.first_tap >= 0 && P.first_tap + P.ntaps <= (long)g.taps.size(),
                          "phase tap range");
                    CHECK(P.ntaps % 4 == 0, "phase taps not a multiple of 4");
                }
                for (const Tap &t : g.taps) {
                    CHECK(t.off >= 0 && t.off <= g.max_off, "tap offset %d", t.off);
                    CHECK(t.off + 7L * g.PP + 7 < tile, "tap outside the LDS tile");
                }
                CHECK(g.tapk.size() == g.taps.size() && g.tapWd.size() == g.taps.size(),
                      "side tables");
            }
        }
    }
    CHECK(tables > 200, "only %ld BIC tables built", tables);

    // union tables of a group of fwhm values against each fwhm's own tables
    const float groups[][5] = {{1.0f, 1.5f, 2.0f, 2.5f, 3.0f}, {0.5f, 3.0f, 0, 0, 0},
                               {2.0f, 1.0f, 1.5f, 0, 0}};
    const int gn[] = {5, 2, 3};
    long compared = 0;
    for (const auto &sz : sizes) {
        for (int M = 1; M <= 6; M++) {
            for (int gi = 0; gi < 3; gi++) {
                const int nf = gn[gi];
                float r[5], tss[5], tso[5];
                for (int f = 0; f < nf; f++) bic_consts(groups[gi][f], r[f], tss[f], tso[f]);
                BicGeometryMulti gm;
                if (!build_bic_geometry_multi(sz[0], sz[1], M, r, tss, tso, nf, gm)) continue;
                for (const TapM &t : gm.taps)
                    CHECK(t.off >= 0 && t.off + 7L * gm.PP + 7 < 4L * gm.PP * gm.PH,
                          "multi tap outside the LDS tile");
                for (int f = 0; f < nf; f++) {
                    BicGeometry g;
                    if (!build_bic_geometry(sz[0], sz[1], M, r[f], tss[f], tso[f], g)) continue;
                    CHECK(g.phases.size() == gm.phases.size(), "phase count");
                    if (g.phases.size() != gm.phases.size()) continue;
                    CHECK(g.xs_log2 == gm.xs_log2, "class width");
                    for (size_t p = 0; p < g.phases.size(); p++) {
                        const Phase &P = g.phases[p];
                        const PhaseM &Q = gm.phases[p];
                        std::vector<int> ks, km;
                        std::vector<float> ws, wm, as, am, bs, bm;
                        for (int k = 0; k < P.ntaps; k++) {
                            const Tap &t = g.taps[P.first_tap + k];
                            if (t.W == 0.0f) continue;
                            ks.push_back(g.tapk[P.first_tap + k]);
                            ws.push_back(t.W);
                            as.push_back(t.a);
                            bs.push_back(t.b);
                        }
                        for (int k = 0; k < Q.ntaps; k++) {
                            const TapM &t = gm.taps[Q.first_tap + k];
                            if (t.W[f] == 0.0f) continue;
                            km.push_back(gm.tapk[Q.first_tap + k]);
                            wm.push_back(t.W[f]);
                            am.push_back(t.a);
                            bm.push_back(t.b);
                        }
                        bool eq = ks == km && ws.size() == wm.size();
                        for (size_t k = 0; eq && k < ws.size(); k++)
                            eq = same_bits(ws[k], wm[k]) && same_bits(as[k], am[k]) &&
                                 same_bits(bs[k], bm[k]);
                        CHECK(eq, "union table != own table: %dx%d M%d fwhm %g phase %zu",
                              sz[0], sz[1], M, (double)groups[gi][f], p);
                        CHECK(same_bits(P.wsum, Q.wsum[f]), "weight sum differs");
                        compared++;
                        CHECK(P.du == Q.du && P.dv == Q.dv && P.sx == Q.sx && P.sy == Q.sy,
                              "bicubic base differs");
                        for (int k = 0; k < 4; k++)
                            CHECK(same_bits(P.wx[k], Q.wx[k]) && same_bits(P.wy[k], Q.wy[k]),
                                  "bicubic weights differ");
                    }
                }
            }
        }
    }
    CHECK(compared > 20000, "only %ld union phases compared", compared);
}

static void check_3d_tables()
{
    long compared = 0;
    const float fwhms[] = {1.0f, 1.5f, 2.5f};
    for (float fwhm : fwhms) {
        float r, tss, tso;
        bic_consts(fwhm, r, tss, tso);
        for (int M = 1; M <= 5; M++) {
            for (int nP = 2; nP <= 9; nP += 3) {
                for (int Mz = 1; Mz <= 6; Mz++) {
                    Geometry3 g;
                    if (!build_3d_geometry(13, 11, nP, M, Mz, r, tss, tso, g)) continue;
                    CHECK(g.nZM == nP * Mz, "planes %d != %d", g.nZM, nP * Mz);
                    CHECK(g.tapv.size() == g.tapo.size() && g.tapk.size() == g.tapv.size(),
                          "3-D side tables");
                    for (const Phase3 &P : g.phases) {
                        CHECK(P.first_tap >= 0 && P.first_tap + P.ntaps <= (long)g.tapv.size(),
                              "3-D phase tap range");
                        CHECK(P.nz == 0 || (P.zlo >= 1 && P.zlo + P.nz <= nP), "3-D planes");
                        CHECK(P.nz <= g.NZ, "3-D plane count over NZ");
                    }
                    for (int o : g.tapo)
                        CHECK(o >= 0 && o < (long)g.NZ * g.PZ, "3-D tap outside the tile");
                    if (Mz > 2) continue;
                    // plane 3z+1 at 3*Mz has plane z's centre at Mz: the same taps
                    Geometry3 g3;
                    if (!build_3d_geometry(13, 11, nP, M, 3 * Mz, r, tss, tso, g3)) continue;
                    const size_t MM = (size_t)M * M;
                    const size_t ncls = (size_t)g.ncls_x * g.ncls_y;
                    CHECK(ncls == (size_t)g3.ncls_x * g3.ncls_y, "3-D classes");
                    for (size_t c = 0; c < ncls; c++) {
                        for (int z = 0; z < g.nZM; z++) {
                            for (size_t s = 0; s < MM; s++) {
                                const Phase3 &A = g.phases[(c * g.nZM + z) * MM + s];
                                const Phase3 &B =
                                    g3.phases[(c * g3.nZM + (3 * z + 1)) * MM + s];
                                bool eq = A.ntaps == B.ntaps && same_bits(A.wsum, B.wsum) &&
                                          A.zlo == B.zlo && A.nz == B.nz;
                                for (int k = 0; eq && k < A.ntaps; k++) {
                                    const float4 a = g.tapv[A.first_tap + k];
                                    const float4 b = g3.tapv[B.first_tap + k];
                                    eq = same_bits(a.x, b.x) && same_bits(a.y, b.y) &&
                                         same_bits(a.z, b.z) && same_bits(a.w, b.w) &&
                                         g.tapk[A.first_tap + k] == g3.tapk[B.first_tap + k];
                                }
                                CHECK(eq, "axial invariant: M%d nP%d Mz%d z%d", M, nP, Mz, z);
                                compared++;
                            }
                        }
                    }
                }
            }
        }
    }
    CHECK(compared > 1000, "only %ld axial phases compared", compared);
}

static void check_plan()
{
    long plans = 0;
    for (int n = 1; n <= 90; n++) {
        for (int nFu = 0; nFu <= n + 2; nFu += (n > 30 ? 7 : 1)) {
            for (int gap = 0; gap <= n + 2; gap += (n > 30 ? 9 : 1)) {
                for (int rolling = 0; rolling <= 1; rolling++) {
                    for (int maxOn = 1; maxOn <= n + 1; maxOn += (n > 30 ? 11 : 1)) {
                        int pl[6];
                        if (esrrf_plan_rolling(n, nFu, gap, rolling, maxOn, pl) != ESRRF_OK) {
                            CHECK(false, "plan rejected n=%d nF=%d gap=%d", n, nFu, gap);
                            continue;
                        }
                        plans++;
                        const int nF = pl[0], g = pl[1], nSR = pl[2], nOn = pl[3], nL = pl[4];
                        CHECK(nF >= 1 && nF <= n, "nF %d of %d", nF, n);
                        CHECK(nOn >= 1 && nOn <= nF, "frames per load %d", nOn);
                        CHECK(nSR >= 1 && (long)(nSR - 1) * g + nF <= n,
                              "SR windows run past the stack: n=%d nF=%d gap=%d nSR=%d", n, nF,
                              g, nSR);
                        for (int r = 1; r <= nSR; r++) {
                            int next = (r - 1) * g + 1, total = 0;
                            for (int l = 0; l < nL; l++) {
                                int first = 0, cnt = 0;
                                CHECK(esrrf_plan_load(pl, r, l, &first, &cnt) == ESRRF_OK,
                                      "load plan");
                                CHECK(first == next && cnt >= 1, "load %d of window %d", l, r);
                                next = first + cnt;
                                total += cnt;
                            }
                            CHECK(total == nF, "window %d loads %d of %d frames", r, total, nF);
                        }
                    }
                }
            }
        }
    }
    CHECK(plans > 10000, "only %ld plans", plans);
}

int main()
{
    check_java_text();
    check_bic_tables();
    check_3d_tables();
    check_plan();
    if (g_fail) {
        fprintf(stderr, "host_check: %d failures\n", g_fail);
        return 1;
    }
    printf("host_check: ok\n");
    return 0;
}
