// Host AddressSanitizer / UndefinedBehaviorSanitizer run of the C ABI
// shims (test infrastructure).
//
// Linked against a mechanism library (libcrs_<mech>.so) and the surrogate
// library (libcrs_mlp.so) whose host code is built with
// -fsanitize=address,undefined (device code unchanged).  It needs no GPU:
// every call below either fails argument validation or has no work
// (n = 0), and both return before any HIP call.  Checked for every entry
// point of include/crs_hip.h and include/crs_mlp.h:
//   - the documented code (CRS_ERR_ARG / CRS_ERR_UNSUPPORTED / CRS_OK) and a
//     message that names the entry point;
//   - messages are per thread (four threads fail at once, each reads its
//     own text back) and per library (a surrogate-library failure leaves
//     the mechanism library's message alone);
//   - no out-of-bounds access, use after free or undefined behaviour in
//     the shim code (the sanitizers abort the run on the first);
//   - the overlap test behind the aliasing checks of crs_dydt / crs_rk4
//     (crs::rows_overlap) agrees with the byte ranges on every small case.
// Prints "ok" and exits 0 when every check passes.
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "crs_hip.h"
#include "crs_mlp.h"
#include "crs_overlap.h"

namespace {

int failures = 0;

void expect(const char *what, int got, int want, const char *(*msg)(void),
            const char *needle)
{
    if (got != want) {
        std::fprintf(stderr, "%s: returned %d, expected %d (%s)\n", what,
                     got, want, msg());
        ++failures;
        return;
    }
    if (want != CRS_OK && needle && !std::strstr(msg(), needle)) {
        std::fprintf(stderr, "%s: message '%s' lacks '%s'\n", what, msg(),
                     needle);
        ++failures;
    }
}

#define MECH(call, want, needle) \
    expect(#call, (call), (want), crs_last_error, (needle))
#define MLP(call, want, needle) \
    expect(#call, (call), (want), crs_mlp_last_error, (needle))

int dummy_exact(const double *, const double *, const double *, double *,
                const int32_t *, const int32_t *, int64_t, int64_t, int,
                void *)
{
    return CRS_OK;
}

void mechanism_library(const crs_mech_info &info)
{
    // host arrays used only as non-NULL pointers: every call below returns
    // before touching them
    alignas(64) static double d[64];
    alignas(64) static int32_t ix[64];
    // a [4 NSP][4] state block and three more of its size, for the overlap
    // checks (addresses only)
    static std::vector<double> rows(16 * (size_t)info.nsp + 64);
    double *p = d;
    double *q = rows.data();
    const int32_t *ip = ix;
    const int64_t big = ((int64_t)1 << 29) + 1;

    MECH(crs_get_mech_info(nullptr), CRS_ERR_ARG, "crs_get_mech_info");

    MECH(crs_eval_conc(p, p, p, p, p, p, p, -1, 4, nullptr), CRS_ERR_ARG,
         "crs_eval_conc: n must be >= 0");
    MECH(crs_eval_conc(p, p, p, p, p, p, p, 4, 3, nullptr), CRS_ERR_ARG,
         "ld must be >= n");
    MECH(crs_eval_conc(nullptr, p, p, p, p, p, p, 4, 4, nullptr),
         CRS_ERR_ARG, "T is NULL");
    MECH(crs_eval_conc(p, p, p, p, p, p, nullptr, 4, 4, nullptr),
         CRS_ERR_ARG, "conc is NULL");
    MECH(crs_eval_conc(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, 0, 0, nullptr), CRS_OK, nullptr);

    MECH(crs_eval_rxn_rates(p, p, nullptr, p, p, 4, 4, nullptr), CRS_ERR_ARG,
         "crs_eval_rxn_rates: C is NULL");
    MECH(crs_eval_rxn_rates(p, p, p, p, nullptr, 4, 4, nullptr),
         info.nrev > 0 ? CRS_ERR_ARG : CRS_OK, "rev is NULL");
    MECH(crs_eval_rxn_rates(nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                            0, nullptr), CRS_OK, nullptr);

    MECH(crs_get_rxn_pres_mod(p, p, p, p, -3, 0, nullptr), CRS_ERR_ARG,
         "crs_get_rxn_pres_mod: n must be >= 0");
    MECH(crs_get_rxn_pres_mod(p, p, p, nullptr, 4, 4, nullptr),
         info.npd > 0 ? CRS_ERR_ARG : CRS_OK, "pres_mod is NULL");
    MECH(crs_get_rxn_pres_mod(nullptr, nullptr, nullptr, nullptr, 0, 0,
                              nullptr), CRS_OK, nullptr);

    MECH(crs_eval_spec_rates(nullptr, p, p, p, p, 4, 4, nullptr),
         CRS_ERR_ARG, "crs_eval_spec_rates: fwd is NULL");
    MECH(crs_eval_spec_rates(p, p, p, p, nullptr, 4, 4, nullptr),
         CRS_ERR_ARG, "dy_N is NULL");
    MECH(crs_eval_spec_rates(nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                             0, nullptr), CRS_OK, nullptr);

    MECH(crs_eval_wdot(p, p, p, p, 4, 4, 7, nullptr), CRS_ERR_ARG,
         "crs_eval_wdot: unknown mode 7");
    MECH(crs_eval_wdot(p, p, p, p, 0, 0, CRS_MODE_MIXED, nullptr),
         (info.modes & (1 << CRS_MODE_MIXED)) ? CRS_OK : CRS_ERR_UNSUPPORTED,
         "not compiled");
    MECH(crs_eval_wdot(p, p, p, nullptr, 4, 4, CRS_MODE_FP64, nullptr),
         CRS_ERR_ARG, "wdot is NULL");
    MECH(crs_eval_wdot(p, p, p, p, 5, 4, CRS_MODE_FP64, nullptr),
         CRS_ERR_ARG, "ld must be >= n");
    MECH(crs_eval_wdot(nullptr, nullptr, nullptr, nullptr, 0, 0,
                       CRS_MODE_FP64, nullptr), CRS_OK, nullptr);

    MECH(crs_eval_wdot_subset(p, p, p, p, ip, -1, 4, CRS_MODE_FP64, nullptr),
         CRS_ERR_ARG, "crs_eval_wdot_subset: m must be >= 0");
    MECH(crs_eval_wdot_subset(p, p, p, p, ip, 1, -1, CRS_MODE_FP64, nullptr),
         CRS_ERR_ARG, "ld must be >= 0");
    MECH(crs_eval_wdot_subset(p, p, p, p, ip, 1, big, CRS_MODE_FP64,
                              nullptr), CRS_ERR_ARG, "exceeds");
    MECH(crs_eval_wdot_subset(p, p, p, p, nullptr, 1, 4, CRS_MODE_FP64,
                              nullptr), CRS_ERR_ARG, "idx is NULL");
    MECH(crs_eval_wdot_subset(p, p, p, p, ip, 1, 4, 9, nullptr), CRS_ERR_ARG,
         "unknown mode");
    MECH(crs_eval_wdot_subset(nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                              0, CRS_MODE_FP64, nullptr), CRS_OK, nullptr);

    MECH(crs_eval_wdot_subset_dev(p, p, p, p, ip, nullptr, 1, 4,
                                  CRS_MODE_FP64, nullptr), CRS_ERR_ARG,
         "crs_eval_wdot_subset_dev: count is NULL");
    MECH(crs_eval_wdot_subset_dev(p, p, p, p, ip, ip, -2, 4, CRS_MODE_FP64,
                                  nullptr), CRS_ERR_ARG, "m must be >= 0");
    MECH(crs_eval_wdot_subset_dev(nullptr, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, 0, 0, CRS_MODE_FP64,
                                  nullptr), CRS_OK, nullptr);

    MECH(crs_dydt(nullptr, nullptr, nullptr, 0, 0, CRS_MODE_MIXED, nullptr),
         (info.modes & (1 << CRS_MODE_MIXED)) ? CRS_OK : CRS_ERR_UNSUPPORTED,
         "crs_dydt: mode 1 not compiled");
    MECH(crs_dydt(p, p, p, 4, 4, 9, nullptr), CRS_ERR_ARG,
         "crs_dydt: unknown mode 9");
    MECH(crs_dydt(p, nullptr, p, 4, 4, CRS_MODE_FP64, nullptr), CRS_ERR_ARG,
         "y is NULL");
    MECH(crs_dydt(nullptr, nullptr, nullptr, 0, 0, CRS_MODE_FP64, nullptr),
         CRS_OK, nullptr);
    // dy written while y and P are read: any shared byte is refused (rows
    // of 4 entries at stride 4: d[0..4 NSP) against d + 1 and d + 4 NSP - 1)
    const int64_t nsp = info.nsp;
    MECH(crs_dydt(p + 40, p, p, 4, 4, CRS_MODE_FP64, nullptr), CRS_ERR_ARG,
         "crs_dydt: y and dy overlap");
    MECH(crs_dydt(p + 40, q, q + 1, 4, 4, CRS_MODE_FP64, nullptr),
         CRS_ERR_ARG, "y and dy overlap");
    MECH(crs_dydt(p + 40, q, q + 4 * nsp - 1, 4, 4, CRS_MODE_FP64, nullptr),
         CRS_ERR_ARG, "y and dy overlap");
    MECH(crs_dydt(q + 4, q + 4 * nsp + 8, q, 4, 4, CRS_MODE_FP64, nullptr),
         CRS_ERR_ARG, "P and dy overlap");

    MECH(crs_rk4(p, p, p, 4, 4, 0.0, 1e-9, -1, nullptr, nullptr, 400.0,
                 nullptr), CRS_ERR_ARG, "crs_rk4: nsteps must be >= 0");
    MECH(crs_rk4(p, p, p, 4, 4, 0.0, 0.0, 1, nullptr, nullptr, 400.0,
                 nullptr), CRS_ERR_ARG, "h must be > 0");
    MECH(crs_rk4(p, p, p, 4, 4, 0.0, std::nan(""), 1, nullptr, nullptr,
                 400.0, nullptr), CRS_ERR_ARG, "h must be > 0");
    MECH(crs_rk4(p, p, p, big, big, 0.0, 1e-9, 1, nullptr, nullptr, 400.0,
                 nullptr), CRS_ERR_ARG, "exceeds");
    MECH(crs_rk4(p, p, p, 4, 4, 0.0, 1e-9, 1, p, nullptr, 400.0, nullptr),
         CRS_ERR_ARG, "T0 and t_ign go together");
    MECH(crs_rk4(p, p, nullptr, 4, 4, 0.0, 1e-9, 1, nullptr, nullptr, 400.0,
                 nullptr), CRS_ERR_ARG, "work is NULL");
    MECH(crs_rk4(p, q, q + 4 * nsp, 4, 4, 0.0, 1e-9, 0, nullptr, nullptr,
                 400.0, nullptr), CRS_OK, nullptr);
    MECH(crs_rk4(p, q, q + 4 * nsp - 2, 4, 4, 0.0, 1e-9, 1, nullptr,
                 nullptr, 400.0, nullptr), CRS_ERR_ARG,
         "crs_rk4: y and work overlap");
    MECH(crs_rk4(q + 4 * nsp, q, q + 4 * nsp, 4, 4, 0.0, 1e-9, 1, nullptr,
                 nullptr, 400.0, nullptr), CRS_ERR_ARG, "P and work overlap");
    MECH(crs_rk4(p, q, q + 4 * nsp, 4, 4, 0.0, 1e-9, 1, p + 8, p + 10, 400.0,
                 nullptr), CRS_ERR_ARG, "t_ign and T0 overlap");
    MECH(crs_rk4(p, q, q + 4 * nsp, 4, 4, 0.0, 1e-9, 1, p + 8, q + 2, 400.0,
                 nullptr), CRS_ERR_ARG, "t_ign and y overlap");
    MECH(crs_rk4(p, q, q + 4 * nsp, 4, 4, 0.0, 1e-9, 1, q + 1, p + 20, 400.0,
                 nullptr), CRS_ERR_ARG, "T0 and y overlap");
    MECH(crs_rk4(nullptr, nullptr, nullptr, 0, 0, 0.0, 1e-9, 3, nullptr,
                 nullptr, 400.0, nullptr), CRS_OK, nullptr);
}

// crs::rows_overlap against the byte ranges themselves: every block shape
// with n <= 3, n <= ld <= 5, up to 3 rows each, at every byte offset of b
// from a within +-4 rows (aligned or not).
void overlap_brute_force()
{
    alignas(64) static char buf[1024];
    const char *a = buf + 512;
    int checked = 0;
    for (int64_t n = 1; n <= 3; ++n)
        for (int64_t ld = n; ld <= 5; ++ld)
            for (int64_t ra = 1; ra <= 3; ++ra)
                for (int64_t rb = 1; rb <= 3; ++rb)
                    for (int64_t d = -32 * ld; d <= 32 * ld; ++d) {
                        bool truth = false;
                        for (int64_t r1 = 0; r1 < ra; ++r1)
                            for (int64_t i1 = 0; i1 < n; ++i1)
                                for (int64_t r2 = 0; r2 < rb; ++r2)
                                    for (int64_t i2 = 0; i2 < n; ++i2) {
                                        const int64_t x = 8 * (r1 * ld + i1);
                                        const int64_t y = d + 8 * (r2 * ld + i2);
                                        if (x < y + 8 && y < x + 8) truth = true;
                                    }
                        const bool got = crs::rows_overlap(a, ra, a + d, rb,
                                                           n, ld);
                        if (got != truth) {
                            std::fprintf(stderr, "rows_overlap(n=%lld ld=%lld "
                                         "ra=%lld rb=%lld d=%lld) = %d, "
                                         "bytes say %d\n", (long long)n,
                                         (long long)ld, (long long)ra,
                                         (long long)rb, (long long)d, got,
                                         truth);
                            ++failures;
                            return;
                        }
                        ++checked;
                    }
    std::printf("rows_overlap: %d cases agree with the byte ranges\n",
                checked);
}

void surrogate_library()
{
    alignas(64) static double d[64];
    alignas(64) static float f[64];
    alignas(64) static int32_t ix[64];
    alignas(64) static unsigned char raw[256];
    double *p = d;
    float *fp = f;
    int32_t *ip = ix;
    void *v = raw;
    const crs_mlp_gate_params gp = {800.0, 2500.0, 1e5, 5e6, 1e-2};

    MLP(crs_mlp_linear_fwd(v, 64, v, 64, fp, 16, 0, 64, 1, v, v, fp, 64, 64,
                           nullptr, 0, nullptr), CRS_ERR_ARG,
        "crs_mlp_linear_fwd: bad sizes");
    MLP(crs_mlp_linear_fwd(v, 64, v, 64, fp, 16, 64, 48, 1, v, v, fp, 64, 64,
                           nullptr, 0, nullptr), CRS_ERR_ARG, "bad sizes");
    MLP(crs_mlp_linear_fwd(v, 64, v, 64, fp, 40, 64, 64, 1, v, v, fp, 64, 64,
                           v, 36, nullptr), CRS_ERR_ARG, "ldat");
    MLP(crs_mlp_linear_fwd(nullptr, 64, v, 64, fp, 16, 64, 64, 1, v, v, fp,
                           64, 64, nullptr, 0, nullptr), CRS_ERR_ARG,
        "NULL input");
    MLP(crs_mlp_linear_fwd(v, 64, v, 64, fp, 16, 64, 64, 1, nullptr, v, fp,
                           64, 64, nullptr, 0, nullptr), CRS_ERR_ARG,
        "NULL activation output");
    MLP(crs_mlp_linear_fwd(v, 64, v, 64, fp, 16, 64, 64, 0, v, v, nullptr,
                           64, 64, nullptr, 0, nullptr), CRS_ERR_ARG,
        "NULL output");
    MLP(crs_mlp_linear_fwd(nullptr, 64, nullptr, 64, nullptr, 0, 64, 64, 1,
                           nullptr, nullptr, nullptr, 64, 64, nullptr, 0,
                           nullptr), CRS_OK, nullptr);

    MLP(crs_mlp_linear_bwd_data(v, 64, v, 64, 16, 64, 40, v, v, 64, nullptr,
                                0, nullptr), CRS_ERR_ARG,
        "crs_mlp_linear_bwd_data: bad sizes");
    MLP(crs_mlp_linear_bwd_data(v, 64, v, 64, 16, 64, 64, nullptr, v, 64,
                                nullptr, 0, nullptr), CRS_ERR_ARG,
        "NULL pointer");
    MLP(crs_mlp_linear_bwd_data(v, 64, v, 64, 40, 64, 64, v, v, 64, v, 33,
                                nullptr), CRS_ERR_ARG, "ldt");
    MLP(crs_mlp_linear_bwd_data(nullptr, 64, nullptr, 64, 0, 64, 64, nullptr,
                                nullptr, 64, nullptr, 0, nullptr), CRS_OK,
        nullptr);

    MLP(crs_mlp_linear_bwd_weight(v, 64, v, 64, 0, 64, 64, fp, fp, nullptr),
        CRS_ERR_ARG, "crs_mlp_linear_bwd_weight: bad sizes");
    MLP(crs_mlp_linear_bwd_weight(v, 32, v, 64, 40, 64, 64, fp, fp, nullptr),
        CRS_ERR_ARG, "leading dimensions");
    MLP(crs_mlp_linear_bwd_weight(v, 64, v, 64, 40, 64, 64, fp, nullptr,
                                  nullptr), CRS_ERR_ARG, "NULL pointer");

    MLP(crs_mlp_transpose_bf16(v, 64, -1, 64, v, 64, nullptr), CRS_ERR_ARG,
        "crs_mlp_transpose_bf16: bad sizes");
    MLP(crs_mlp_transpose_bf16(v, 32, 4, 64, v, 64, nullptr), CRS_ERR_ARG,
        "bad sizes");
    MLP(crs_mlp_transpose_bf16(nullptr, 64, 4, 64, v, 64, nullptr),
        CRS_ERR_ARG, "NULL pointer");
    MLP(crs_mlp_transpose_bf16(nullptr, 64, 0, 64, nullptr, 64, nullptr),
        CRS_OK, nullptr);

    MLP(crs_mlp_features(p, p, p, nullptr, 8, 8, 13, fp, fp, v, 14, nullptr),
        CRS_ERR_ARG, "crs_mlp_features: bad sizes");
    MLP(crs_mlp_features(p, p, p, nullptr, 8, 4, 13, fp, fp, v, 64, nullptr),
        CRS_ERR_ARG, "bad sizes");
    MLP(crs_mlp_features(p, p, p, ip, 8, 4, 13, fp, nullptr, v, 64, nullptr),
        CRS_ERR_ARG, "NULL pointer");
    MLP(crs_mlp_features(nullptr, nullptr, nullptr, nullptr, 0, 0, 13,
                         nullptr, nullptr, nullptr, 64, nullptr), CRS_OK,
        nullptr);

    MLP(crs_mlp_targets(p, 8, 4, 13, p, fp, fp, fp, 13, nullptr),
        CRS_ERR_ARG, "crs_mlp_targets: bad sizes");
    MLP(crs_mlp_targets(p, 8, 8, 13, p, fp, fp, fp, 12, nullptr),
        CRS_ERR_ARG, "bad sizes");
    MLP(crs_mlp_targets(p, 8, 8, 13, nullptr, fp, fp, fp, 13, nullptr),
        CRS_ERR_ARG, "NULL pointer");
    MLP(crs_mlp_targets(nullptr, 0, 0, 13, nullptr, nullptr, nullptr,
                        nullptr, 13, nullptr), CRS_OK, nullptr);

    MLP(crs_mlp_decode(fp, 12, nullptr, 8, 8, 13, p, fp, fp, p, nullptr),
        CRS_ERR_ARG, "crs_mlp_decode: bad sizes");
    MLP(crs_mlp_decode(fp, 64, nullptr, 8, 4, 13, p, fp, fp, p, nullptr),
        CRS_ERR_ARG, "bad sizes");
    MLP(crs_mlp_decode(fp, 64, ip, 8, 4, 13, p, fp, fp, nullptr, nullptr),
        CRS_ERR_ARG, "NULL pointer");
    MLP(crs_mlp_decode(nullptr, 64, nullptr, 0, 0, 13, nullptr, nullptr,
                       nullptr, nullptr, nullptr), CRS_OK, nullptr);

    MLP(crs_mlp_mse_grad(fp, fp, 0, 13, 64, 64, v, 64, fp, nullptr),
        CRS_ERR_ARG, "crs_mlp_mse_grad: bad sizes");
    MLP(crs_mlp_mse_grad(fp, fp, 8, 13, 12, 64, v, 64, fp, nullptr),
        CRS_ERR_ARG, "bad sizes");
    MLP(crs_mlp_mse_grad(fp, nullptr, 8, 13, 64, 64, v, 64, fp, nullptr),
        CRS_ERR_ARG, "NULL pointer");

    MLP(crs_mlp_adam(fp, fp, fp, fp, v, 16, 1e-3f, 0.9f, 0.999f, 1e-8f, 0,
                     0.f, nullptr), CRS_ERR_ARG, "crs_mlp_adam: bad sizes");
    MLP(crs_mlp_adam(fp, fp, fp, fp, nullptr, 16, 1e-3f, 0.9f, 0.999f, 1e-8f,
                     1, 0.f, nullptr), CRS_ERR_ARG, "NULL pointer");

    MLP(crs_mlp_to_bf16(fp, v, -1, nullptr), CRS_ERR_ARG,
        "crs_mlp_to_bf16: n must be >= 0");
    MLP(crs_mlp_to_bf16(fp, nullptr, 8, nullptr), CRS_ERR_ARG,
        "NULL pointer");
    MLP(crs_mlp_to_bf16(nullptr, nullptr, 0, nullptr), CRS_OK, nullptr);

    MLP(crs_mlp_gate(p, p, p, p, 8, 4, 13, gp, p, 3, p, ip, ip, ip, nullptr),
        CRS_ERR_ARG, "crs_mlp_gate: bad sizes");
    MLP(crs_mlp_gate(p, p, p, p, 8, 8, 13, gp, nullptr, 3, p, ip, ip, ip,
                     nullptr), CRS_ERR_ARG, "NULL pointer");
    MLP(crs_mlp_gate(nullptr, nullptr, nullptr, nullptr, 0, 0, 13, gp,
                     nullptr, 0, nullptr, nullptr, nullptr, nullptr,
                     nullptr), CRS_OK, nullptr);

    MLP(crs_mlp_pack_tiles(v, 15, 64, 0, v, nullptr), CRS_ERR_ARG,
        "crs_mlp_pack_tiles: bad shape");
    MLP(crs_mlp_pack_tiles(nullptr, 16, 64, 0, v, nullptr), CRS_ERR_ARG,
        "NULL pointer");

    // crs_mlp_infer: shape, sizes, pointers, alignment
    MLP(crs_mlp_infer(p, p, p, 8, 8, 13, 32, 256, 3, 64, v, fp, fp, fp, p,
                      fp, fp, p, gp, p, 3, p, ip, ip, ip, nullptr),
        CRS_ERR_UNSUPPORTED, "crs_mlp_infer: fused kernel is built for");
    MLP(crs_mlp_infer(p, p, p, 8, 8, 1, 64, 256, 3, 64, v, fp, fp, fp, p, fp,
                      fp, p, gp, p, 3, p, ip, ip, ip, nullptr), CRS_ERR_ARG,
        "bad sizes");
    MLP(crs_mlp_infer(p, p, p, 8, 8, 13, 64, 256, 3, 64, v, fp, fp, fp, p,
                      fp, fp, p, gp, p, 9, p, ip, ip, ip, nullptr),
        CRS_ERR_ARG, "n_elem=9");
    MLP(crs_mlp_infer(p, p, p, 8, 8, 13, 64, 256, 3, 64, v, fp, fp, fp, p,
                      fp, fp, p, gp, p, 3, p, nullptr, ip, ip, nullptr),
        CRS_ERR_ARG, "NULL pointer");
    MLP(crs_mlp_infer(p, p, p, 8, 8, 13, 64, 256, 3, 64, raw + 8, fp, fp, fp,
                      p, fp, fp, p, gp, p, 3, p, ip, ip, ip, nullptr),
        CRS_ERR_ARG, "alignment");
    MLP(crs_mlp_infer(nullptr, nullptr, nullptr, 0, 0, 13, 64, 256, 3, 64,
                      nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                      nullptr, nullptr, gp, nullptr, 0, nullptr, nullptr,
                      nullptr, nullptr, nullptr), CRS_OK, nullptr);

    crs_mlp_net net;
    std::memset(&net, 0, sizeof net);
    net.nsp = 13;
    net.kin = 64;
    net.hidden = 256;
    net.layers = 3;
    net.kout = 64;
    net.gate = gp;
    MLP(crs_mlp_eval_gated(nullptr, p, p, p, p, 8, 8, dummy_exact, 0, ip, ip,
                           ip, nullptr), CRS_ERR_ARG,
        "crs_mlp_eval_gated: NULL network");
    MLP(crs_mlp_eval_gated(&net, p, p, p, p, 8, 8, nullptr, 0, ip, ip, ip,
                           nullptr), CRS_ERR_ARG, "exact function");
    MLP(crs_mlp_eval_gated(&net, p, p, p, p, 8, 4, dummy_exact, 0, ip, ip,
                           ip, nullptr), CRS_ERR_ARG, "bad sizes");
    MLP(crs_mlp_eval_gated(&net, p, p, p, p, 8, 8, dummy_exact, 0, ip,
                           nullptr, ip, nullptr), CRS_ERR_ARG,
        "NULL fallback list");
    MLP(crs_mlp_eval_gated(&net, nullptr, nullptr, nullptr, nullptr, 0, 0,
                           dummy_exact, 0, nullptr, nullptr, nullptr,
                           nullptr), CRS_OK, nullptr);
    MLP(crs_mlp_eval_gated_noalloc(&net, p, p, p, p, ((int64_t)1 << 22) + 1,
                                   ((int64_t)1 << 22) + 1, dummy_exact, 0,
                                   nullptr), CRS_ERR_ARG,
        "crs_mlp_eval_gated_noalloc: n = 4194305 exceeds");
    MLP(crs_mlp_eval_gated_noalloc(&net, p, p, p, p, -1, 0, dummy_exact, 0,
                                   nullptr), CRS_ERR_ARG, "bad sizes");
    MLP(crs_mlp_eval_gated_noalloc(&net, nullptr, nullptr, nullptr, nullptr,
                                   0, 0, dummy_exact, 0, nullptr), CRS_OK,
        nullptr);
}

// Each thread fails with its own n and reads its own message back after
// every thread has failed.
void per_thread_messages()
{
    const int nthreads = 4;
    std::atomic<int> arrived{0};
    std::vector<int> ok(nthreads, 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([t, &arrived, &ok] {
            const int rc = crs_eval_conc(nullptr, nullptr, nullptr, nullptr,
                                         nullptr, nullptr, nullptr,
                                         -(int64_t)(t + 1), 0, nullptr);
            arrived.fetch_add(1);
            while (arrived.load() < nthreads)
                std::this_thread::yield();
            const std::string want =
                "(got " + std::to_string(-(t + 1)) + ")";
            ok[t] = rc == CRS_ERR_ARG &&
                    std::strstr(crs_last_error(), want.c_str()) != nullptr;
        });
    for (auto &th : pool)
        th.join();
    for (int t = 0; t < nthreads; ++t)
        if (!ok[t]) {
            std::fprintf(stderr, "thread %d did not read its own message\n",
                         t);
            ++failures;
        }
}

// A failure in the surrogate library leaves the mechanism library's
// message in place, and the other way round.
void per_library_messages()
{
    crs_eval_conc(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                  nullptr, -7, 0, nullptr);
    crs_mlp_to_bf16(nullptr, nullptr, -5, nullptr);
    if (!std::strstr(crs_last_error(), "(got -7)") ||
        !std::strstr(crs_mlp_last_error(), "crs_mlp_to_bf16")) {
        std::fprintf(stderr, "messages shared between libraries: '%s' / "
                     "'%s'\n", crs_last_error(), crs_mlp_last_error());
        ++failures;
    }
}

}  // namespace

int main()
{
    crs_mech_info info;
    std::memset(&info, 0, sizeof info);
    if (crs_get_mech_info(&info) != CRS_OK || info.nsp < 2 || !info.name) {
        std::fprintf(stderr, "crs_get_mech_info failed\n");
        return 1;
    }
    std::printf("library %s: nsp %d nr %d nrev %d npd %d modes %d\n",
                info.name, info.nsp, info.nr, info.nrev, info.npd,
                info.modes);
    overlap_brute_force();
    mechanism_library(info);
    surrogate_library();
    per_thread_messages();
    per_library_messages();
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
