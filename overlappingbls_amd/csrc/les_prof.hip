// Per-family device time of one step (les_profile_step): the launches of
// each instrumented kernel family are bracketed by HIP events on the stream
// they run on, so the measurement sees the families as the step runs them,
// including the kernels that share the GPU with the auxiliary stream.
#include <vector>
#include "les_internal.hpp"

namespace les {

namespace {
const char* const NAMES[FAM_N] = {
    "xline.deriv", "yline.deriv", "xline.filter", "yline.filter", "xline.pad", "yline.pad", "xline.unpad",
    "yline.unpad", "x_r2c", "x_c2r", "yline.spec", "tdma", "vort_pad_x", "cross_unpad_x", "scalar_unpad_x",
    "lag_contract_x", "rhs_ab2", "gradp_project", "sgs_stress", "scalar_flux", "ddz", "divstress", "cfl_max",
    "plane_sums", "press_prep", "lagrange_point", "theta_point"};
}  // namespace

const char* fam_name(int f) { return (f >= 0 && f < FAM_N) ? NAMES[f] : nullptr; }

struct FamProf {
    struct Rec {
        int fam;
        double bytes;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    size_t used = 0;
    bool overflow = false;
    int only = -1;  // the one family timed, or -1 for every family
};
FamProf* g_prof = nullptr;

FamScope::FamScope(hipStream_t st, int fam, double bytes) {
    if (!g_prof || (g_prof->only >= 0 && g_prof->only != fam)) return;
    if (g_prof->used >= g_prof->recs.size()) {
        g_prof->overflow = true;
        return;
    }
    slot = (int)g_prof->used++;
    s = st;
    FamProf::Rec& r = g_prof->recs[slot];
    r.fam = fam;
    r.bytes = bytes;
    (void)hipEventRecord(r.a, s);
}

FamScope::~FamScope() {
    if (slot >= 0 && g_prof) (void)hipEventRecord(g_prof->recs[slot].b, s);
}

}  // namespace les

using namespace les;

extern "C" {

int les_nfamilies(void) { return FAM_N; }
const char* les_family_name(int fam) { return fam_name(fam); }

int les_profile_step(les_ctx* c, int jt, double* ms, double* bytes, int* launches) {
    return les_profile_run(c, jt, 1, -1, ms, bytes, launches);
}

int les_profile_run(les_ctx* c, int jt0, int nsteps, int family, double* ms, double* bytes, int* launches) {
    if (!c || !ms || !bytes || !launches || g_prof || nsteps < 1 || family < -1 || family >= FAM_N)
        return LES_E_BADARG;
    FamProf prof;
    prof.only = family;
    // a step launches fewer than 512 instrumented kernels
    prof.recs.resize((size_t)512 * (size_t)nsteps);
    size_t made = 0;
    hipError_t e = hipSuccess;
    for (; made < prof.recs.size() && e == hipSuccess; ++made) {
        e = hipEventCreate(&prof.recs[made].a);
        if (e == hipSuccess) e = hipEventCreate(&prof.recs[made].b);
    }
    int st = LES_E_HIP;
    if (e == hipSuccess) {
        int failed = 0;
        g_prof = &prof;
        st = les_run(c, jt0, nsteps, &failed);
        g_prof = nullptr;
    }
    for (int f = 0; f < FAM_N; ++f) {
        ms[f] = 0.0;
        bytes[f] = 0.0;
        launches[f] = 0;
    }
    if (st == LES_OK) {
        for (size_t i = 0; i < prof.used; ++i) {
            float t = 0.0f;
            const FamProf::Rec& r = prof.recs[i];
            if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) {
                st = LES_E_HIP;
                break;
            }
            ms[r.fam] += t;
            bytes[r.fam] += r.bytes;
            launches[r.fam] += 1;
        }
        if (prof.overflow) st = LES_E_BADARG;
    }
    for (size_t i = 0; i < made; ++i) {
        (void)hipEventDestroy(prof.recs[i].a);
        (void)hipEventDestroy(prof.recs[i].b);
    }
    return st;
}

}  // extern "C"
