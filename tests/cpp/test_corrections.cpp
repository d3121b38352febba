// The shape of the reference's tests/corrections_test.cpp: cable (and
// geometric) corrections on the GPU compared with the CPU restatement of
// src/corrections.cpp.  The reference compares with FITS fixtures bit for bit
// (tests/common.hpp:56); its fixtures are not available, and the GPU path
// reduces the angle in FP64 and evaluates sin/cos in FP32, so the bar here is
// |dV| <= 2e-6 |V| against the FP64-maths CPU result.
#include "blink_oracle.h"
#include "common.hpp"
#include "pacer_imager_mi355x.h"

static blink_ctx* g_ctx = nullptr;

static void ok(int rc) {
    if (rc != BLINK_OK) throw TestFailed(blink_last_error(g_ctx));
}

static double worst_rel(const MemoryBuffer<std::complex<float>>& a, const std::vector<float>& b) {
    double worst = 0;
    for (size_t i = 0; i < a.size(); i++) {
        const std::complex<double> g(a[i]);
        const std::complex<double> r(b[2 * i], b[2 * i + 1]);
        worst = std::max(worst, std::abs(g - r) / std::max(std::abs(r), 1e-30));
    }
    return worst;
}

void test_corrections(bool geometric) {
    const int n_ant = 128, n_ch = 4, n_t = 2, n_bl = n_ant * (n_ant + 1) / 2;
    SplitMix64 g(geometric ? 7 : 8);
    std::vector<double> x(n_ant), y(n_ant), z(n_ant), cable(n_ant), freq(n_ch);
    for (int a = 0; a < n_ant; a++) {
        x[a] = 300 * g.normal();
        y[a] = 300 * g.normal();
        z[a] = 5 * g.normal();
        cable[a] = -1722 + 1946 * g.uniform();  // the MWA metafits EL_ range
    }
    for (int ch = 0; ch < n_ch; ch++) freq[ch] = (169.6 + 0.04 * ch) * 1e6;
    std::vector<float> u(n_bl), v(n_bl), w(n_bl);
    oracle_zenith_uvw(n_ant, x.data(), y.data(), z.data(), u.data(), v.data(), w.data());
    const size_t n = (size_t)n_t * n_ch * n_bl * 4;
    MemoryBuffer<std::complex<float>> vis(n);
    std::vector<float> ref(2 * n);
    for (size_t i = 0; i < n; i++) {
        vis[i] = std::complex<float>((float)g.normal(), (float)g.normal());
        ref[2 * i] = vis[i].real();
        ref[2 * i + 1] = vis[i].imag();
    }
    if (geometric) oracle_geometric_corrections(ref.data(), n_ant, n_ch, n_t, w.data(), freq.data());
    else oracle_cable_corrections(ref.data(), n_ant, n_ch, n_t, cable.data(), freq.data());
    vis.to_gpu();
    MemoryBuffer<double> f(n_ch), c(n_ant);
    MemoryBuffer<float> wd(n_bl);
    for (int i = 0; i < n_ch; i++) f[i] = freq[i];
    for (int i = 0; i < n_ant; i++) c[i] = cable[i];
    for (int i = 0; i < n_bl; i++) wd[i] = w[i];
    f.to_gpu();
    c.to_gpu();
    wd.to_gpu();
    if (geometric)
        ok(blink_apply_geometric_corrections(g_ctx, (float*)vis.data(), n_ant, n_ch, n_t, wd.data(),
                                             f.data(), nullptr));
    else
        ok(blink_apply_cable_corrections(g_ctx, (float*)vis.data(), n_ant, n_ch, n_t, c.data(),
                                         f.data(), nullptr));
    vis.to_cpu();
    const double err = worst_rel(vis, ref);
    std::cout << (geometric ? "geometric" : "cable") << " correction: max |dV|/|V| = " << err << std::endl;
    if (!(err < 2e-6)) throw TestFailed("corrections differ from the CPU reference");
}

void test_cable_lengths_correction_gpu() { test_corrections(false); }
void test_geometric_correction_gpu() { test_corrections(true); }

int main() {
    if (blink_create(0, &g_ctx) != BLINK_OK) {
        std::cerr << "no GPU" << std::endl;
        return 1;
    }
    int rc = 0;
    try {
        RUN_TEST(test_cable_lengths_correction_gpu);
        RUN_TEST(test_geometric_correction_gpu);
        std::cout << "All tests passed." << std::endl;
    } catch (std::exception& ex) {
        std::cerr << ex.what() << std::endl;
        rc = 1;
    }
    blink_destroy(g_ctx);
    return rc;
}
