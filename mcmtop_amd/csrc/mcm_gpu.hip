// mcm_gpu.hip -- host runtime and C ABI of the MI355X MCM SPH time step.
//
// One context = one run on one GPU: device-resident structure-of-arrays
// state, one HIP stream, a captured two-step hipGraph for the fused path,
// and the granular stage entry points that replace the reference's
// argument-less step subroutines (MCM/mcm_solution.f90:41-116).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <string>
#include <vector>

#include "../../include/mcm_gpu.h"
#include "mcm_wave.hip"
#include "mcm_output.h"
#include "mcm_textdeck.h"

namespace {

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                         \
            return MCM_GPU_ERR_HIP;                                                               \
        }                                                                                         \
    } while (0)

// Deck content.  The readers fill it with the values as written (material
// cards before the set-up routines, 1-based material ids); finish_deck then
// applies the reference's defaults and material set-up.
struct HostDeck {
    int np = 0, nummat = 0, disctype = 0, h_opt = 0, tcrit_opt = 0, contacttype = 0, krtype = 1;
    int maxnbr = 0, massopt = 0, init_h_opt = 0, init_rhoe = 0, mirr_opt = 0;
    int veloc_opt = 0;                     // 1: XSPH
    int ncontmats = -1, cont_opt = 0;      // penalty contact cards (K and n in MatParams)
    double drelax_scale = 0.0;             // dynamic relaxation when 0.1 < scale < 1
    int nthp[3] = {0, 0, 0};               // base acceleration on where 1
    double base_a[3] = {0.0, 0.0, 0.0};
    int sym_flag[6] = {0, 0, 0, 0, 0, 0};
    double tssf = 0, itss = 0, endtime = 0, sym_x[6] = {0, 0, 0, 0, 0, 0};
    std::vector<MatParams> mat;
    std::vector<MatParams> mat_raw;              // material cards before the set-up routines (defaults applied)
    std::vector<double> mat_mass, mat_rho_min;   // material mass and rho_min cards
    std::vector<double> x, v, pm, ph, pr, pe;
    std::vector<int> pmat, dispbc;
};

}  // namespace

struct mcm_gpu_batch;

struct mcm_gpu_ctx {
    HostDeck deck;
    DevState d{};
    Scalars* sc = nullptr;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph2 = nullptr;   // two fused steps: parity 0 then 1
    int cur = 0;                       // parity of the current h / rho buffers
    int phase = 0;                     // 0: step start, 1: after strain (x in x1, h/rho in next)
    bool fresh = false;                // strain records hold this step's search (set by the search and
                                       // the initialisation, cleared by a move or a fused step)
    unsigned done = 0;                 // stages of the current step already run (ST_* bits)
    int nb = 0;                        // blocks over np
    int corr_one_pass = 0;             // corrected SPH: one-pass moments instead of the reference's three passes
    int pair_mapping = 0;              // 0 thread per particle, 1 wavefront per particle (lists up to 128)
    std::vector<double> xzero;
    std::vector<void*> allocs;
    std::string err;
    int device = 0;
    void* out_buf = nullptr;           // staging buffer of the packed output fields (first use)
    bool own_stream = true;            // false for a run of a batch (the batch's stream)
    mcm_gpu_batch* batch = nullptr;    // the batch this run belongs to, if any
    std::vector<mcm_gpu_adjoint*> recs; // recordings of this run (released with it)
};

namespace {

// ------------------------------------------------------------------- deck
constexpr int32_t kDeckMagic = 0x4D434D31;
constexpr int32_t kContactTag = 0x434F4E54;   // 'CONT': penalty contact section after the particles
constexpr int32_t kOptsTag = 0x4F505453;      // 'OPTS': dynamic relaxation and base acceleration section

template <typename T>
bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

// material constants as mcm_matin.f90:46-56 / mcm_sets{1,3,10}.f90 /
// mcm_printm.f90:84-100 leave them
void setup_material(MatParams& m) {
    if (m.visc_type == 0) m.visc_type = 1;
    if (m.av_l == 0.0) m.av_l = 0.06;
    if (m.av_q == 0.0) m.av_q = 1.5;
    double* cm = m.strinput;
#define C(a, b) cm[((a) - 1) + 5 * ((b) - 1)]
    if (m.model == 1) {                       // mcm_sets1.f90:23-51
        double prop6 = C(2, 1), prop7 = C(3, 1);
        for (int i = 2; i <= 5; i++)
            for (int j = 1; j <= 5; j++) C(i, j) = 0.0;
        double q1 = C(1, 1) * C(1, 2) / ((1.0 + C(1, 2)) * (1.0 - 2.0 * C(1, 2)));
        double q2 = C(1, 1) * 0.5 / (1.0 + C(1, 2));
        double q3 = q1 + 2.0 * q2;
        C(2, 1) = q3; C(3, 2) = q3; C(4, 3) = q3; C(5, 4) = q2;
        C(2, 2) = q1; C(2, 3) = q1; C(3, 1) = q1; C(3, 3) = q1; C(4, 1) = q1; C(4, 2) = q1;
        C(2, 6) = C(1, 6);
        C(1, 6) = C(1, 1) * C(1, 4) / (C(1, 1) - C(1, 4));
        C(1, 7) = prop6;
        C(2, 7) = prop7;
    } else if (m.model == 3) {                // mcm_sets3.f90:25-43
        double prop7 = C(2, 6);
        for (int i = 2; i <= 5; i++)
            for (int j = 1; j <= 6; j++) C(i, j) = 0.0;
        double q1 = C(1, 1) * C(1, 2) / ((1.0 + C(1, 2)) * (1.0 - 2.0 * C(1, 2)));
        double q2 = C(1, 1) * 0.5 / (1.0 + C(1, 2));
        double q3 = q1 + 2.0 * q2;
        C(2, 1) = q3;
        C(5, 4) = q2;
        C(2, 4) = C(1, 1) * C(1, 4) / (C(1, 1) - C(1, 4));
        C(2, 6) = prop7;
        m.efps_flag = (int)lround(C(1, 6));
    } else if (m.model == 10) {               // mcm_sets10.f90:23-36
        double prop[48];
        for (int i = 0; i < 48; i++) prop[i] = cm[i];
        if (prop[3] == 0.0) prop[3] = (double)(-1.e30f);
        if (prop[6] == 0.0) prop[6] = (double)(1.000000001f);
        for (int i = 0; i < 48; i++) cm[i] = prop[i];
        for (int i = 2; i <= 16; i++)
            if (prop[i + 14] < prop[i + 15]) cm[15] = (double)i;
        m.ispall = (int)cm[6];
    }
#undef C
    if ((m.eos == 4 || m.eos == 41) && m.eosinput[7] == 0.0) m.eosinput[7] = 1.0;
    if ((m.eos == 13 || m.eos == 28) && m.eosinput[3] == 0.0) m.eosinput[3] = 1.0;
    // mcm_calcg.f90:31-59
    if (m.model == 1 || m.model == 3) m.g = m.strinput[0] / (2.0 * (1.0 + m.strinput[5]));
    else if (m.model == 4) m.g = m.strinput[5];
    else if (m.model == 9) m.g = 0.0;
    else if (m.model == 10) m.g = m.strinput[0];
}

// Defaults, material set-up and checks shared by both deck formats (mass
// option 0 shares the material mass card equally).
int finish_deck(mcm_gpu_ctx* ctx) {
    HostDeck& D = ctx->deck;
    const std::vector<double>& mat_mass = D.mat_mass;
    // control defaults, mcm_control_v2.f90:193-196
    if (D.tssf == 0.0) D.tssf = 0.8;
    if (D.krtype == 0) D.krtype = 1;
    if (D.maxnbr == 0) D.maxnbr = 40;
    D.mat_raw.clear();
    for (int m = 0; m < D.nummat; m++) {
        MatParams& M = D.mat[m];
        // material card defaults, mcm_matin.f90:46-56
        if (M.visc_type == 0) M.visc_type = 1;
        if (M.av_l == 0.0) M.av_l = 0.06;
        if (M.av_q == 0.0) M.av_q = 1.5;
        if (M.rho_max == 0.0) M.rho_max = 1.0E+10;
        M.pad_ = 0;
        M.efps_flag = 0; M.ispall = 0;
        if (M.model == 9 || M.model == 10) {
            if (M.eos != 1 && M.eos != 4 && M.eos != 13 && M.eos != 28 && M.eos != 41) {
                ctx->err = "unsupported equation of state"; return MCM_GPU_ERR_MODEL;
            }
        } else if (M.model == 1 || M.model == 3 || M.model == 4) {
            if (M.eos != 0) { ctx->err = "models 1/3/4 take no EOS (mcm_matin.f90:72-79)"; return MCM_GPU_ERR_MODEL; }
            // the cap model's mode tests are complete for soil (1) and rock (2)
            // only; for another ltype the reference keeps the last call's mode
            const int lt = (int)M.strinput[1 + 5 * 5];
            if (M.model == 4 && lt != 1 && lt != 2) {
                ctx->err = "cap model (material model 4): ltype (card 2 field 6) must be 1 (soil) or 2 (rock)";
                return MCM_GPU_ERR_MODEL;
            }
        } else {
            ctx->err = "material model not supported on the GPU path (1, 3, 4, 9, 10 are)";
            return MCM_GPU_ERR_MODEL;
        }
        D.mat_raw.push_back(M);
        if (M.model == 4 && cap_sets4(M.strinput)) {   // mcm_matin.f90:152-161 -> mcm_sets4
            ctx->err = "cap model (material model 4): the initial cap iteration of mcm_sets4 does not converge";
            return MCM_GPU_ERR_MODEL;
        }
        setup_material(M);
        if (M.visc_type != 1) { ctx->err = "only nodal viscosity (visc_type 1) exists"; return MCM_GPU_ERR_MODEL; }
    }
    const long n = D.np;
    for (long i = 0; i < n; i++) {
        D.pmat[i] -= 1;
        if (D.pmat[i] < 0 || D.pmat[i] >= D.nummat) { ctx->err = "particle with bad material id"; return MCM_GPU_ERR_ARG; }
    }
    // mcm_calcmass.f90:122-137 (massopt 0: equal share of the material mass)
    if (D.massopt == 0) {
        std::vector<int> cnt(D.nummat, 0);
        for (long i = 0; i < n; i++) cnt[D.pmat[i]]++;
        for (long i = 0; i < n; i++) D.pm[i] = mat_mass[D.pmat[i]] / (double)cnt[D.pmat[i]];
    }
    // mcm_contact_in.f90:40-46 sets K and n only for "every material pair"
    if (D.contacttype > 0 && D.ncontmats != -1) {
        ctx->err = "penalty contact needs ncontmats = -1 (the reference implements no other contact definition)";
        return MCM_GPU_ERR_ARG;
    }
    D.veloc_opt = D.veloc_opt == 1 ? 1 : 0;   // mcm_update_velocity.f90:61 tests veloc_opt == 1
    if (D.krtype != 1 && D.krtype != 0) { ctx->err = "only the cubic B-spline kernel exists"; return MCM_GPU_ERR_ARG; }
    if (D.disctype != 0 && D.disctype != 1) { ctx->err = "disctype must be 0 or 1"; return MCM_GPU_ERR_ARG; }
    return MCM_GPU_OK;
}

// binary deck written by mcmtop_amd/deck.py (Deck.write_bin)
int read_binary_deck(mcm_gpu_ctx* ctx, FILE* f) {
    HostDeck& D = ctx->deck;
    int32_t magic, ints[19];
    double reals[9];
    bool ok = rd(f, &magic, 1) && magic == kDeckMagic && rd(f, ints, 19) && rd(f, reals, 9);
    if (!ok) { ctx->err = "not an MCM binary deck"; return MCM_GPU_ERR_IO; }
    D.np = ints[0]; D.nummat = ints[1]; D.disctype = ints[2]; D.h_opt = ints[3]; D.tcrit_opt = ints[4];
    D.contacttype = ints[5]; D.krtype = ints[6]; D.maxnbr = ints[7]; D.massopt = ints[8];
    D.init_h_opt = ints[9]; D.init_rhoe = ints[10]; D.mirr_opt = ints[11];
    D.veloc_opt = ints[18];
    for (int k = 0; k < 6; k++) D.sym_flag[k] = ints[12 + k];
    D.tssf = reals[0]; D.itss = reals[1]; D.endtime = reals[2];
    for (int k = 0; k < 6; k++) D.sym_x[k] = reals[3 + k];
    if (D.np <= 0 || D.nummat <= 0 || D.nummat > MCM_MAX_MAT) {
        ctx->err = "deck: bad particle or material count"; return MCM_GPU_ERR_ARG;
    }
    D.mat.assign(D.nummat, MatParams{});
    D.mat_mass.assign(D.nummat, 0.0);
    D.mat_rho_min.assign(D.nummat, 0.0);
    for (int m = 0; m < D.nummat; m++) {
        int32_t mi[3];
        double mr[7];
        MatParams& M = D.mat[m];
        if (!(rd(f, mi, 3) && rd(f, mr, 7) && rd(f, M.strinput, 48) && rd(f, M.eosinput, 24))) {
            ctx->err = "deck truncated in material cards"; return MCM_GPU_ERR_IO;
        }
        M.model = mi[0]; M.eos = mi[1]; M.visc_type = mi[2];
        M.rho = mr[0]; M.av_q = mr[1]; M.av_l = mr[2]; M.h = mr[4];
        M.rho_max = mr[6];
        D.mat_mass[m] = mr[3];
        D.mat_rho_min[m] = mr[5];
    }
    long n = D.np;
    D.x.resize(3 * n); D.v.resize(3 * n); D.pmat.resize(n); D.dispbc.resize(n);
    D.pm.resize(n); D.ph.resize(n); D.pr.resize(n); D.pe.resize(n);
    ok = rd(f, D.x.data(), 3 * n) && rd(f, D.v.data(), 3 * n) && rd(f, D.pmat.data(), n) &&
         rd(f, D.dispbc.data(), n) && rd(f, D.pm.data(), n) && rd(f, D.ph.data(), n) &&
         rd(f, D.pr.data(), n) && rd(f, D.pe.data(), n);
    if (!ok) { ctx->err = "deck truncated in particle data"; return MCM_GPU_ERR_IO; }
    if (D.contacttype > 0) {   // penalty contact section: tag, ncontmats, cont_opt, K and n per material
        int32_t ct[3];
        if (!(rd(f, ct, 3) && ct[0] == kContactTag)) { ctx->err = "contact deck without its contact section"; return MCM_GPU_ERR_IO; }
        D.ncontmats = ct[1];
        D.cont_opt = ct[2];
        for (int m = 0; m < D.nummat; m++) {
            double kn[2];
            if (!rd(f, kn, 2)) { ctx->err = "deck truncated in contact data"; return MCM_GPU_ERR_IO; }
            D.mat[m].k_cont = kn[0];
            D.mat[m].n_cont = kn[1];
        }
    }
    int32_t ot[4];
    if (rd(f, ot, 4) && ot[0] == kOptsTag) {   // dynamic relaxation and base acceleration
        double ov[4];
        if (!rd(f, ov, 4)) { ctx->err = "deck truncated in its options section"; return MCM_GPU_ERR_IO; }
        for (int k = 0; k < 3; k++) { D.nthp[k] = ot[1 + k]; D.base_a[k] = ov[1 + k]; }
        D.drelax_scale = ov[0];
    }
    return MCM_GPU_OK;
}

// the reference's version-2 text deck (mcm_getinput.f90:89-127)
int read_text_deck(mcm_gpu_ctx* ctx, FILE* f) {
    using textdeck::Fmt;
    HostDeck& D = ctx->deck;
    textdeck::Cards C;
    C.f = f;
    std::string card;
    auto bad = [&](const std::string& what, int code = MCM_GPU_ERR_IO) {
        ctx->err = "text deck, line " + std::to_string(C.line) + ": " + what;
        return code;
    };
    auto next = [&](const char* what) { return C.next(card) ? 0 : bad(std::string("missing ") + what); };
    int rc;
    // title and format version, (a78, i2)
    if ((rc = next("title card"))) return rc;
    { Fmt F(card); F.skip(78); long vern = F.i(2); if (!F.ok || vern != 2) return bad("only input format version 2 is read"); }
    // control cards (mcm_control_v2.f90:86-124)
    long axopt, thnode, ntrans, init_v_opt, veloc_opt, nlcur, nthp[3];
    double drelax;
    if ((rc = next("control card 1"))) return rc;
    { Fmt F(card); axopt = F.i(5); D.disctype = F.i(5); D.nummat = F.i(10); D.np = F.i(10);
      F.i(10); F.i(10); F.i(10); if (!F.ok) return bad("control card 1"); }
    if ((rc = next("control card 2"))) return rc;
    { Fmt F(card); D.endtime = F.e(10); D.tssf = F.e(10); D.itss = F.e(10); drelax = F.e(10);
      if (!F.ok) return bad("control card 2"); }
    if ((rc = next("control card 3"))) return rc;
    { Fmt F(card); F.e(10); F.i(5); F.e(10); thnode = F.i(10); ntrans = F.i(5); F.i(10); F.i(5); F.i(5);
      if (!F.ok) return bad("control card 3"); }
    if ((rc = next("control card 4"))) return rc;
    { Fmt F(card); F.i(5); init_v_opt = F.i(5); D.massopt = F.i(5); D.init_h_opt = F.i(5); D.init_rhoe = F.i(5);
      if (!F.ok) return bad("control card 4"); }
    if ((rc = next("control card 5"))) return rc;
    { Fmt F(card); D.contacttype = F.i(5); D.tcrit_opt = F.i(5); D.veloc_opt = veloc_opt = F.i(5); D.mirr_opt = F.i(5);
      nlcur = F.i(5); for (int k = 0; k < 3; k++) nthp[k] = F.i(5); if (!F.ok) return bad("control card 5"); }
    if ((rc = next("control card 6"))) return rc;
    { Fmt F(card); D.h_opt = F.i(5); D.krtype = F.i(5); D.maxnbr = F.i(5); if (!F.ok) return bad("control card 6"); }
    if ((rc = next("control card 7"))) return rc;
    if (axopt != 3) return bad("only 3-D problems (axopt 3) run on the GPU path", MCM_GPU_ERR_ARG);
    D.drelax_scale = drelax;
    for (int k = 0; k < 3; k++) D.nthp[k] = (int)nthp[k];
    (void)nlcur;   // load curves are read by nothing on the step's path
    if (D.np <= 0 || D.nummat <= 0 || D.nummat > MCM_MAX_MAT) return bad("bad particle or material count", MCM_GPU_ERR_ARG);
    // material cards (mcm_matin.f90:38-203)
    D.mat.assign(D.nummat, MatParams{});
    D.mat_mass.assign(D.nummat, 0.0);
    D.mat_rho_min.assign(D.nummat, 0.0);
    std::vector<char> seen(D.nummat, 0);
    for (int m = 0; m < D.nummat; m++) {
        if ((rc = next("material control card"))) return rc;
        long n, model, eos, visc;
        double rho, av_q, av_l;
        { Fmt F(card); n = F.i(5); model = F.i(5); rho = F.e(10); eos = F.i(5); F.i(5); F.e(10); visc = F.i(5);
          av_q = F.e(10); av_l = F.e(10); if (!F.ok) return bad("material control card"); }
        if (n < 1 || n > D.nummat || seen[n - 1]) return bad("material number out of range", MCM_GPU_ERR_ARG);
        seen[n - 1] = 1;
        MatParams& M = D.mat[n - 1];
        M.model = (int)model; M.eos = (int)eos; M.rho = rho;
        M.visc_type = (int)visc; M.av_l = av_l; M.av_q = av_q;
        if ((rc = next("material options card"))) return rc;
        { Fmt F(card); D.mat_mass[n - 1] = F.e(10); M.h = F.e(10); D.mat_rho_min[n - 1] = F.e(10);
          M.rho_max = F.e(10); if (!F.ok) return bad("material options card"); }
        if ((rc = next("material title card"))) return rc;
        int per_card;
        if (model == 1 || model == 3 || model == 4 || model == 9) per_card = 5;
        else if (model == 10) per_card = 8;
        else return bad("material model not implemented", MCM_GPU_ERR_MODEL);
        for (int j = 0; j < 6; j++) {
            if ((rc = next("material property card"))) return rc;
            Fmt F(card);
            for (int k = 0; k < per_card; k++) M.strinput[per_card * j + k] = F.e(10);
            if (!F.ok) return bad("material property card");
        }
        if (eos != 0) {
            if ((rc = next("eos title card"))) return rc;
            // values per card: eos 1 reads 8 then 1 (mcm_matin.f90:177-185)
            std::vector<int> cards;
            if (eos == 1) cards = {8, 1};
            else if (eos == 4 || eos == 41) cards = {8};
            else if (eos == 13) cards = {5};
            else if (eos == 28) cards = {2};
            else return bad("eos type not implemented", MCM_GPU_ERR_MODEL);
            int at = 0;
            for (int cnt : cards) {
                if ((rc = next("eos property card"))) return rc;
                Fmt F(card);
                for (int k = 0; k < cnt; k++) M.eosinput[at + k] = F.e(10);
                if (!F.ok) return bad("eos property card");
                at += cnt;
            }
        }
    }
    // particle cards (mcm_nodein.f90:34-125), (i8, i5, 3e20.0, i7)
    const long np = D.np;
    D.x.assign(3 * np, 0.0); D.v.assign(3 * np, 0.0); D.pmat.assign(np, 0); D.dispbc.assign(np, 0);
    D.pm.assign(np, 0.0); D.ph.assign(np, 0.0); D.pr.assign(np, 0.0); D.pe.assign(np, 0.0);
    for (long j = 0; j < np; j++) {
        if ((rc = next("particle card"))) return rc;
        Fmt F(card);
        long id = F.i(8);
        D.dispbc[j] = (int)F.i(5);
        for (int k = 0; k < 3; k++) D.x[3 * j + k] = F.e(20);
        D.pmat[j] = (int)F.i(7);
        if (!F.ok) return bad("particle card");
        if (id != j + 1) return bad("particle card out of order", MCM_GPU_ERR_ARG);
        if (D.pmat[j] == 0) return bad("particle with zero material", MCM_GPU_ERR_ARG);
    }
    // additional particle data, (i8, 4e20.0) on an 80-column card
    if (D.massopt == 1 || D.init_h_opt == 1 || D.init_rhoe == 1) {
        for (long j = 0; j < np; j++) {
            if ((rc = next("additional particle card"))) return rc;
            Fmt F(card);
            long id = F.i(8);
            double v4[4];
            for (int k = 0; k < 4; k++) v4[k] = F.e(20);
            if (!F.ok) return bad("additional particle card");
            if (id != j + 1) return bad("additional particle card out of order", MCM_GPU_ERR_ARG);
            if (D.massopt == 1) D.pm[j] = v4[0];
            if (D.init_h_opt == 1) D.ph[j] = v4[1];
            if (D.init_rhoe == 1) { D.pr[j] = v4[2]; D.pe[j] = v4[3]; }
        }
    }
    // time-history particles (10 ids a card) and pressure transducers: output
    // only, read past
    for (long k = 0; k < (thnode + 9) / 10; k++)
        if ((rc = next("time history card"))) return rc;
    for (long k = 0; k < ntrans; k++)
        if ((rc = next("transducer card"))) return rc;
    // initial velocities with the generation rule (mcm_Inputvel.f90:29-129),
    // (i8, 3e10.0)
    if (init_v_opt == 1) {
        auto read_v = [&](long* nid, double* v) {
            Fmt F(card);
            *nid = F.i(8);
            for (int k = 0; k < 3; k++) v[k] = F.e(10);
            return F.ok;
        };
        long nid;
        if ((rc = next("velocity card"))) return rc;
        if (!read_v(&nid, &D.v[0])) return bad("velocity card");
        if (nid != 1) return bad("first velocity card is not for particle 1", MCM_GPU_ERR_ARG);
        bool gen = false;
        for (long i = 1; i < np; i++) {
            double* vi = &D.v[3 * i];
            const double* vp = &D.v[3 * (i - 1)];
            if (!gen) {
                if ((rc = next("velocity card"))) return rc;
                if (!read_v(&nid, vi)) return bad("velocity card");
                if (nid > i + 1) {
                    if (nid > np) return bad("velocity card id exceeds the particle count", MCM_GPU_ERR_ARG);
                    for (int k = 0; k < 3; k++)
                        if (vi[k] != vp[k]) return bad("velocity for generation is not constant", MCM_GPU_ERR_ARG);
                    gen = true;
                } else if (nid < i + 1) {
                    return bad("velocity card id less than the previous card's", MCM_GPU_ERR_ARG);
                }
            } else {
                for (int k = 0; k < 3; k++) vi[k] = vp[k];
                if (i + 1 == nid) gen = false;
            }
        }
    }
    // boundary planes (mcm_inp_symmetry_v2.f90:23-98), (6i10) and (6e10.0)
    for (int k = 0; k < 6; k++) { D.sym_flag[k] = 0; D.sym_x[k] = 0.0; }
    if (D.mirr_opt == 1) {
        if ((rc = next("boundary plane flag card"))) return rc;
        { Fmt F(card); for (int k = 0; k < 6; k++) D.sym_flag[k] = (int)F.i(10); if (!F.ok) return bad("boundary plane flag card"); }
        if ((rc = next("boundary plane coordinate card"))) return rc;
        { Fmt F(card); for (int k = 0; k < 6; k++) D.sym_x[k] = F.e(10); if (!F.ok) return bad("boundary plane coordinate card"); }
    }
    // penalty contact cards (mcm_contact_in.f90:22-45): (2i5), then K and n
    // per material as (2E10.5) when every material pair is in contact
    if (D.contacttype > 0) {
        if ((rc = next("contact card"))) return rc;
        { Fmt F(card); D.ncontmats = (int)F.i(5); D.cont_opt = (int)F.i(5); if (!F.ok) return bad("contact card"); }
        if (D.ncontmats == -1)
            for (int m = 0; m < D.nummat; m++) {
                if ((rc = next("contact material card"))) return rc;
                Fmt F(card);
                D.mat[m].k_cont = F.ed(10, 5);
                D.mat[m].n_cont = F.ed(10, 5);
                if (!F.ok) return bad("contact material card");
            }
    }
    // base acceleration card (mcm_baseaccel_in.f90:25-31): (3e20.0)
    if (D.nthp[0] == 1 || D.nthp[1] == 1 || D.nthp[2] == 1) {
        if ((rc = next("base acceleration card"))) return rc;
        Fmt F(card);
        for (int k = 0; k < 3; k++) D.base_a[k] = F.e(20);
        if (!F.ok) return bad("base acceleration card");
    }
    return MCM_GPU_OK;
}

// A deck is the binary format when it starts with its magic number, else the
// reference's text format.  Values as written (no defaults applied).
int read_deck_raw(mcm_gpu_ctx* ctx, const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) { ctx->err = std::string("cannot open deck ") + path; return MCM_GPU_ERR_IO; }
    int32_t magic = 0;
    const bool binary = fread(&magic, 4, 1, f) == 1 && magic == kDeckMagic;
    rewind(f);
    const int rc = binary ? read_binary_deck(ctx, f) : read_text_deck(ctx, f);
    fclose(f);
    return rc;
}

int read_deck(mcm_gpu_ctx* ctx, const char* path) {
    int rc = read_deck_raw(ctx, path);
    return rc ? rc : finish_deck(ctx);
}

// the binary deck layout of Deck.write_bin (mcmtop_amd/deck.py)
int write_binary_deck(const HostDeck& D, const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) return MCM_GPU_ERR_IO;
    const int32_t magic = kDeckMagic;
    int32_t ints[19] = {D.np, D.nummat, D.disctype, D.h_opt, D.tcrit_opt, D.contacttype, D.krtype, D.maxnbr,
                        D.massopt, D.init_h_opt, D.init_rhoe, D.mirr_opt, 0, 0, 0, 0, 0, 0, D.veloc_opt};
    for (int k = 0; k < 6; k++) ints[12 + k] = D.sym_flag[k];
    double reals[9] = {D.tssf, D.itss, D.endtime, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 6; k++) reals[3 + k] = D.sym_x[k];
    bool ok = fwrite(&magic, 4, 1, f) == 1 && fwrite(ints, 4, 19, f) == 19 && fwrite(reals, 8, 9, f) == 9;
    for (int m = 0; m < D.nummat && ok; m++) {
        const MatParams& M = D.mat[m];
        int32_t mi[3] = {M.model, M.eos, M.visc_type};
        double mr[7] = {M.rho, M.av_q, M.av_l, D.mat_mass[m], M.h, D.mat_rho_min[m], M.rho_max};
        ok = fwrite(mi, 4, 3, f) == 3 && fwrite(mr, 8, 7, f) == 7 && fwrite(M.strinput, 8, 48, f) == 48 &&
             fwrite(M.eosinput, 8, 24, f) == 24;
    }
    const size_t n = D.np;
    ok = ok && fwrite(D.x.data(), 8, 3 * n, f) == 3 * n && fwrite(D.v.data(), 8, 3 * n, f) == 3 * n &&
         fwrite(D.pmat.data(), 4, n, f) == n && fwrite(D.dispbc.data(), 4, n, f) == n &&
         fwrite(D.pm.data(), 8, n, f) == n && fwrite(D.ph.data(), 8, n, f) == n && fwrite(D.pr.data(), 8, n, f) == n &&
         fwrite(D.pe.data(), 8, n, f) == n;
    if (ok && D.contacttype > 0) {
        const int32_t ct[3] = {kContactTag, D.ncontmats, D.cont_opt};
        ok = fwrite(ct, 4, 3, f) == 3;
        for (int m = 0; m < D.nummat && ok; m++) {
            const double kn[2] = {D.mat[m].k_cont, D.mat[m].n_cont};
            ok = fwrite(kn, 8, 2, f) == 2;
        }
    }
    if (ok && (D.drelax_scale != 0.0 || D.nthp[0] || D.nthp[1] || D.nthp[2])) {
        const int32_t ot[4] = {kOptsTag, D.nthp[0], D.nthp[1], D.nthp[2]};
        const double ov[4] = {D.drelax_scale, D.base_a[0], D.base_a[1], D.base_a[2]};
        ok = fwrite(ot, 4, 4, f) == 4 && fwrite(ov, 8, 4, f) == 4;
    }
    ok = fclose(f) == 0 && ok;
    return ok ? MCM_GPU_OK : MCM_GPU_ERR_IO;
}

// ------------------------------------------------------------- allocation
template <typename T>
int dalloc(mcm_gpu_ctx* ctx, T** p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&q, n * sizeof(T)));
    HIPCHK(hipMemsetAsync(q, 0, n * sizeof(T), ctx->stream));
    ctx->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

void dfree(mcm_gpu_ctx* ctx, void* p) {
    if (!p) return;
    auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), p);
    if (it != ctx->allocs.end()) ctx->allocs.erase(it);
    (void)hipFree(p);
}

// grow an array keeping its first keep elements
template <typename T>
int dregrow(mcm_gpu_ctx* ctx, T** p, size_t n_new, size_t keep) {
    T* q = nullptr;
    int rc = dalloc(ctx, &q, n_new);
    if (rc) return rc;
    if (keep) HIPCHK(hipMemcpyAsync(q, *p, keep * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
    dfree(ctx, *p);
    *p = q;
    return 0;
}

int alloc_neighbour_arrays(mcm_gpu_ctx* ctx) {
    DevState& d = ctx->d;
    int rc = 0;
    rc |= dalloc(ctx, &d.nbr, (size_t)d.ncap * d.npad);
    rc |= dalloc(ctx, &d.cell_count, d.cellcap);
    rc |= dalloc(ctx, &d.cell_start, d.cellcap);
    for (int k = 0; k < 3; k++) {
        rc |= dalloc(ctx, &d.gstart[k], d.cellcap);
        rc |= dalloc(ctx, &d.gcount[k], d.cellcap);
    }
    rc |= dalloc(ctx, &d.scan_tmp, d.cellcap / SCAN_TILE + 2);
    rc |= dalloc(ctx, &d.gbase, d.cellcap);
    rc |= dalloc(ctx, &d.gany, d.cellcap);
    if (d.contacttype > 0) rc |= dalloc(ctx, &d.cont, (size_t)d.ncap * d.npad);
    return rc;
}

void free_neighbour_arrays(mcm_gpu_ctx* ctx) {
    DevState& d = ctx->d;
    if (d.cont) { dfree(ctx, d.cont); d.cont = nullptr; }
    dfree(ctx, d.nbr); dfree(ctx, d.cell_count); dfree(ctx, d.cell_start); dfree(ctx, d.scan_tmp); dfree(ctx, d.gbase); dfree(ctx, d.gany);
    for (int k = 0; k < 3; k++) { dfree(ctx, d.gstart[k]); dfree(ctx, d.gcount[k]); }
}

int alloc_all(mcm_gpu_ctx* ctx) {
    DevState& d = ctx->d;
    const size_t n = d.np, u = (size_t)d.np + d.ngcap;
    int rc = 0;
    for (int k = 0; k < 3; k++) {
        rc |= dalloc(ctx, &d.x0[k], u); rc |= dalloc(ctx, &d.x1[k], u); rc |= dalloc(ctx, &d.v[k], u);
    }
    rc |= dalloc(ctx, &d.mass, u);
    for (int k = 0; k < 2; k++) { rc |= dalloc(ctx, &d.hbuf[k], u); rc |= dalloc(ctx, &d.rbuf[k], u); }
    rc |= dalloc(ctx, &d.pmat, u);
    rc |= dalloc(ctx, &d.rs, 2 * RS_SLOTS * u);
    rc |= dalloc(ctx, &d.rm, 2 * RM_SLOTS * u);
    rc |= dalloc(ctx, &d.gpar, d.ngcap);
    rc |= dalloc(ctx, &d.gsgn, d.ngcap);
    rc |= dalloc(ctx, &d.dispbc, n);
    double** per[] = {&d.hold, &d.h0, &d.rho0, &d.p, &d.e, &d.c, &d.vabs, &d.pcut, &d.p_cut, &d.efps, &d.fail,
                      &d.q, &d.mindist, &d.tracerod, &d.einc, &d.etry, &d.rhoold, &d.critts_p, &d.defps_prev};
    for (auto pp : per) rc |= dalloc(ctx, pp, n);
    for (int k = 0; k < 6; k++) {
        rc |= dalloc(ctx, &d.sig[k], n); rc |= dalloc(ctx, &d.sd[k], n); rc |= dalloc(ctx, &d.alfa[k], n);
        rc |= dalloc(ctx, &d.rod[k], n);
    }
    for (int k = 0; k < 3; k++) { rc |= dalloc(ctx, &d.spin[k], n); rc |= dalloc(ctx, &d.a[k], n); }
    rc |= dalloc(ctx, &d.yield, n);
    if (d.asym) {   // the cap model's state and the stress's lower triangle
        rc |= dalloc(ctx, &d.capa, n);
        rc |= dalloc(ctx, &d.ptype, n);
        for (int k = 0; k < 3; k++) rc |= dalloc(ctx, &d.sigl[k], n);
        rc |= dalloc(ctx, &d.rml, 4 * u);
    }
    rc |= dalloc(ctx, &d.cell_of, n); rc |= dalloc(ctx, &d.slot_of, n); rc |= dalloc(ctx, &d.sorted, n);
    rc |= dalloc(ctx, &d.unsorted, n);
    rc |= dalloc(ctx, &d.cand, 4 * n);
    rc |= dalloc(ctx, &d.cand_id, 2 * n);
    rc |= dalloc(ctx, &d.spos, n);
    rc |= dalloc(ctx, &d.nreal, n); rc |= dalloc(ctx, &d.ntot, n);
    if (d.skin > 0.0) {
        d.skin_cap = d.ncap + 64;
        d.skin_blkcap = std::max(4096, d.cellcap / 8);
        rc |= dalloc(ctx, &d.skin_list, (size_t)d.skin_cap * d.npad);
        rc |= dalloc(ctx, &d.skin_n, n);
        rc |= dalloc(ctx, &d.skin_pos, n);
        rc |= dalloc(ctx, &d.skin_id, n);
        rc |= dalloc(ctx, &d.skin_blk, n);
        rc |= dalloc(ctx, &d.skin_mat, n);
        // workgroups of k_skin_blocks, one partial each
        d.skin_nred = std::max(1, std::min(128, (d.skin_blkcap + MCM_BLOCK - 1) / MCM_BLOCK));
        rc |= dalloc(ctx, &d.skin_red, 8 * (size_t)d.skin_nred);
        rc |= dalloc(ctx, &d.skin_base, 4 * n);
        rc |= dalloc(ctx, &d.xh, 4 * n);
        for (int k = 0; k < 2; k++) rc |= dalloc(ctx, &d.skin_bstat[k], 8 * (size_t)d.skin_blkcap);
    }
    if (d.contacttype > 0) {
        rc |= dalloc(ctx, &d.ncont, n);
        for (int k = 0; k < 3; k++) rc |= dalloc(ctx, &d.rep[k], n);
    }
    if (d.veloc_opt == 1)
        for (int k = 0; k < 3; k++) rc |= dalloc(ctx, &d.sv[k], n);
    rc |= dalloc(ctx, &d.red, (size_t)ctx->nb * RED_C + 64);
    rc |= dalloc(ctx, &d.mat, d.nummat);
    rc |= dalloc(ctx, &d.sc, 1);
    ctx->sc = d.sc;
    rc |= dalloc(ctx, &d.range, 1);
    if (!rc) HIPCHK(hipMemsetAsync(d.range, 0xff, sizeof(int), ctx->stream));   // no fast path before a search
    rc |= alloc_neighbour_arrays(ctx);
    return rc;
}

// ----------------------------------------------------------------- launches
int blocks_for(long n, int bs) { return (int)((n + bs - 1) / bs); }

// Launch sizes of the grid-stride kernels over cells and ghosts: bounded by
// the capacity and by the particle count (a run holds at most a few cells
// and ghosts per particle), so that the launches of a batch of small runs
// stay small.
int cell_blocks(const DevState& d) {
    return std::max(1, std::min({blocks_for(d.cellcap, MCM_BLOCK), 4096, blocks_for(d.np, MCM_BLOCK)}));
}
int tile_blocks(const DevState& d) {
    return std::max(1, std::min(blocks_for(d.cellcap, SCAN_TILE), blocks_for(d.np, SCAN_TILE) + 1));
}
int skin_block_blocks(const DevState& d) { return d.skin_nred; }
int ghost_blocks(const DevState& d) {
    return std::max(1, std::min({blocks_for(d.ngcap, MCM_BLOCK), 1024, blocks_for(d.np, 2048)}));
}

// 0 plain SPH, 1 corrected in the reference's pass order, 2 corrected in one pass
int corr_variant(const mcm_gpu_ctx* ctx) { return ctx->d.disctype == 1 ? (ctx->corr_one_pass ? 2 : 1) : 0; }

// the wavefront-per-particle pair kernels hold up to 2 x 64 list entries; runs
// with penalty contact or XSPH take the thread-per-particle kernels
bool wave_capable(const mcm_gpu_ctx* ctx) { return ctx->d.contacttype == 0 && ctx->d.veloc_opt == 0; }
bool use_wave(const mcm_gpu_ctx* ctx) {
    return ctx->pair_mapping == 1 && wave_capable(ctx) && ctx->d.ncap <= 64 * WAVE_SLOTS && corr_variant(ctx) != 2 &&
           !ctx->d.asym;
}

// list-build kernel of a run: 0 one material, 1 several, 2 wavefront per
// particle; with a neighbour skin 3 one material, 4 several
int list_variant(const mcm_gpu_ctx* c) {
    if (c->d.skin > 0.0) return c->d.nummat > 1 ? 4 : 3;
    return c->pair_mapping == 1 && wave_capable(c) ? 2 : (c->d.nummat > 1 ? 1 : 0);
}

// strain-pass kernel of a run: corr_variant, plus 3 for the wavefront mapping
int pair_variant(const mcm_gpu_ctx* c) { return corr_variant(c) + (use_wave(c) ? 3 : 0); }

// momentum kernel of a run: as the strain pass, or 5 + corr_variant when
// XSPH moves the particles after the velocity update
// (runs with a cap-model material: 8 + corr_variant, or 11 + corr_variant
// with XSPH, the kernels that read the stress's lower triangle)
// constitutive kernel of the fused step: 1 when every material is f3dm10 (the
// kernel compiled for that model alone), else 0
int const_variant(const mcm_gpu_ctx* c) {
    for (int m = 0; m < c->d.nummat; m++)
        if (c->deck.mat[m].model != 10) return 0;
    return 1;
}

int mom_variant(const mcm_gpu_ctx* c) {
    if (c->d.asym) return (c->d.veloc_opt == 1 ? 11 : 8) + corr_variant(c);
    return c->d.veloc_opt == 1 ? 5 + corr_variant(c) : pair_variant(c);
}

// What a launch of the step spans, per run.  kind_blocks gives the
// workgroups one run needs in a launch of that kind and kernel variant v;
// 0 means the run takes no part in it.
enum LaunchKind { L_ONE, L_PART, L_CELLS, L_TILES, L_GSCAN, L_GFILL, L_GSTRESS, L_LIST, L_PAIR, L_MOM, L_CFOLD,
                  L_REP, L_XSPH, L_SKIN, L_SKINP, L_SKINB, L_CONST, L_NKIND };
constexpr int kNVariant = 16;
static_assert(kNVariant > MCM_MAX_MAT && kNVariant > 13, "launch variants");

int kind_blocks(const mcm_gpu_ctx* c, int kind, int v) {
    const DevState& d = c->d;
    switch (kind) {
        case L_ONE: return v == 0 ? 1 : 0;
        case L_PART: return v == 0 ? c->nb : 0;
        case L_CELLS: return v == 0 ? cell_blocks(d) : 0;
        case L_TILES: return v == 0 ? tile_blocks(d) : 0;
        case L_GSCAN: return v < d.nplanes ? 1 : 0;
        case L_GFILL: return v < d.nplanes ? std::min(ghost_blocks(d), 512) : 0;
        case L_GSTRESS: return v == 0 && d.boundary ? ghost_blocks(d) : 0;
        case L_LIST: return list_variant(c) == v ? c->nb : 0;
        case L_PAIR: return pair_variant(c) == v ? c->nb : 0;
        case L_MOM: return mom_variant(c) == v ? c->nb : 0;
        case L_CFOLD: return d.nummat == v ? 1 : 0;
        case L_REP: return v == 0 && d.contacttype > 0 ? c->nb : 0;
        case L_XSPH: return v == 0 && d.veloc_opt == 1 ? c->nb : 0;
        case L_SKIN: return v == 0 && d.skin > 0.0 ? 1 : 0;
        case L_SKINP: return v == 0 && d.skin > 0.0 ? c->nb : 0;
        case L_SKINB: return v == 0 && d.skin > 0.0 ? skin_block_blocks(d) : 0;
        case L_CONST: return const_variant(c) == v ? c->nb : 0;
    }
    return 0;
}

}  // namespace

// Several independent runs stepped together: every launch of the step
// covers all of them, with a block map from workgroup to (run, block).
struct mcm_gpu_batch {
    std::vector<mcm_gpu_ctx*> inst;
    hipStream_t stream = nullptr;
    int device = 0;
    mcm_gpu_options opt{};               // options of every run
    DevState* states = nullptr;          // device copy of every run's DevState
    BlockEntry* entries = nullptr;       // the block maps of all launch kinds
    size_t entries_cap = 0;
    int map_off[L_NKIND][kNVariant] = {};
    int map_n[L_NKIND][kNVariant] = {};
    int max_planes = 0;
    hipGraphExec_t graph2 = nullptr;     // two batched steps: parity 0 then 1
    int cur = 0;
    bool stale = true;                   // states and maps must be uploaded again
    std::string err;
};

namespace {

// Where a launch goes: one run, or every run of a batch.
struct Target {
    mcm_gpu_ctx* ctx;        // one run (null for a batch)
    mcm_gpu_batch* batch;    // a batch (null for one run)
    hipStream_t stream;
    int planes() const { return batch ? batch->max_planes : ctx->d.nplanes; }
};

Target one(mcm_gpu_ctx* c) { return Target{c, nullptr, c->stream}; }
Target all(mcm_gpu_batch* b) { return Target{nullptr, b, b->stream}; }

// Launch kernel kern(d, args...) with threads per workgroup over every run of
// the target that takes part in (kind, v).  A batched launch passes a state
// that carries only the block map and the runs' states; the kernel's enter()
// replaces it by its run's state.
template <typename... P, typename... A>
void launch(const Target& T, int kind, int v, int threads, void (*kern)(DevState, P...), A... args) {
    if (T.batch == nullptr) {
        const int n = kind_blocks(T.ctx, kind, v);
        if (n > 0) kern<<<n, threads, 0, T.stream>>>(T.ctx->d, args...);
        return;
    }
    const int n = T.batch->map_n[kind][v];
    if (n == 0) return;
    DevState d{};
    d.bmap = T.batch->entries + T.batch->map_off[kind][v];
    d.states = T.batch->states;
    kern<<<n, threads, 0, T.stream>>>(d, args...);
}

// the search up to the lists: grid, counting sort, candidate records, ghosts
void launch_search(const Target& T, int cur) {
    launch(T, L_ONE, 0, 1, k_grid_setup);
    launch(T, L_CELLS, 0, MCM_BLOCK, k_clear_cells);
    launch(T, L_PART, 0, MCM_BLOCK, k_bin, cur);
    launch(T, L_TILES, 0, 1024, k_scan_tiles);
    launch(T, L_ONE, 0, 1024, k_scan_top);
    launch(T, L_TILES, 0, 1024, k_scan_apply);
    launch(T, L_PART, 0, MCM_BLOCK, k_scatter);
    launch(T, L_PART, 0, MCM_BLOCK, k_sort_cells);
    launch(T, L_PART, 0, MCM_BLOCK, k_fill_cand, cur);
    for (int p = 0; p < T.planes(); p++) {
        launch(T, L_GSCAN, p, 1024, k_ghost_scan, p);
        launch(T, L_GFILL, p, MCM_BLOCK, k_ghost_fill, p, cur);
    }
}

// the list build
void launch_lists(const Target& T, int cur) {
    launch(T, L_LIST, 0, MCM_BLOCK, k_list<false, false>, cur);
    launch(T, L_LIST, 1, MCM_BLOCK, k_list<true, false>, cur);
    launch(T, L_LIST, 2, MCM_BLOCK, k_list_wave, cur);
    // neighbour skin: displacement since the build, the check, the build
    // where needed, then the filter
    launch(T, L_SKINP, 0, MCM_BLOCK, k_skin_stats, cur);
    launch(T, L_SKINB, 0, MCM_BLOCK, k_skin_blocks, cur);
    launch(T, L_SKIN, 0, 64, k_skin_check);
    launch(T, L_LIST, 3, MCM_BLOCK, k_list<false, true>, cur);
    launch(T, L_LIST, 3, MCM_BLOCK, k_skin_filter<false>, cur);
    launch(T, L_LIST, 4, MCM_BLOCK, k_list<true, true>, cur);
    launch(T, L_LIST, 4, MCM_BLOCK, k_skin_filter<true>, cur);
}

void launch_neighbours(const Target& T, int cur) {
    launch_search(T, cur);
    launch_lists(T, cur);
}

// energies (and, after a move, coordinate bounds) from the block partials
void launch_c_fold(const Target& T, int with_bounds) {
    launch(T, L_CFOLD, 1, MCM_BLOCK, k_c_fold<1>, with_bounds);
    launch(T, L_CFOLD, 2, MCM_BLOCK, k_c_fold<2>, with_bounds);
    launch(T, L_CFOLD, 3, MCM_BLOCK, k_c_fold<3>, with_bounds);
    launch(T, L_CFOLD, 4, MCM_BLOCK, k_c_fold<4>, with_bounds);
    launch(T, L_CFOLD, 5, MCM_BLOCK, k_c_fold<5>, with_bounds);
    launch(T, L_CFOLD, 6, MCM_BLOCK, k_c_fold<6>, with_bounds);
    launch(T, L_CFOLD, 7, MCM_BLOCK, k_c_fold<7>, with_bounds);
    launch(T, L_CFOLD, 8, MCM_BLOCK, k_c_fold<8>, with_bounds);
}

// strain pass of the fused step: stages rod, spin, tracerod and x1
void launch_strain(const Target& T, int cur) {
    launch(T, L_PAIR, 0, MCM_BLOCK, k_strain_const<MODE_FSTRAIN, 0, false>, cur);
    launch(T, L_PAIR, 1, MCM_BLOCK, k_strain_const<MODE_FSTRAIN, 1, false>, cur);
    launch(T, L_PAIR, 2, MCM_BLOCK, k_strain_const<MODE_FSTRAIN, 2, false>, cur);
    launch(T, L_PAIR, 3, MCM_BLOCK, k_strain_wave<0>, cur);
    launch(T, L_PAIR, 4, MCM_BLOCK, k_strain_wave<1>, cur);
}

// update_h + constitutive + critts of the fused step, then dt and ptime, then
// the contact repulsion of runs with penalty contact (mcm_solution.f90:58-64)
void launch_constitutive(const Target& T, int cur) {
    launch(T, L_CONST, 0, MCM_BLOCK, k_strain_const<MODE_FCONST, 0, false>, cur);
    launch(T, L_CONST, 1, MCM_BLOCK, k_strain_const<MODE_FCONST, 0, false, 1 << 10>, cur);
    launch(T, L_ONE, 0, MCM_BLOCK, k_strain_fold<MODE_FCONST>);
    launch(T, L_REP, 0, MCM_BLOCK, k_rep_cont, cur);
}

void launch_ghost_stress(const Target& T, int cur) { launch(T, L_GSTRESS, 0, MCM_BLOCK, k_ghost_stress, cur); }

// momentum + boundcond + energies + velocity + move of the fused step (with
// XSPH: up to the velocity update, then the smoothing pass and the move)
void launch_momentum_kernel(const Target& T, int cur) {
    launch(T, L_MOM, 0, MCM_BLOCK, k_momentum<MODE_FUSED, 0, false>, cur);
    launch(T, L_MOM, 1, MCM_BLOCK, k_momentum<MODE_FUSED, 1, false>, cur);
    launch(T, L_MOM, 2, MCM_BLOCK, k_momentum<MODE_FUSED, 2, false>, cur);
    launch(T, L_MOM, 3, MCM_BLOCK, k_momentum_wave<0>, cur);
    launch(T, L_MOM, 4, MCM_BLOCK, k_momentum_wave<1>, cur);
    launch(T, L_MOM, 5, MCM_BLOCK, k_momentum<MODE_FVEL, 0, false>, cur);
    launch(T, L_MOM, 6, MCM_BLOCK, k_momentum<MODE_FVEL, 1, false>, cur);
    launch(T, L_MOM, 7, MCM_BLOCK, k_momentum<MODE_FVEL, 2, false>, cur);
    launch(T, L_MOM, 8, MCM_BLOCK, k_momentum<MODE_FUSED, 4, false>, cur);
    launch(T, L_MOM, 9, MCM_BLOCK, k_momentum<MODE_FUSED, 5, false>, cur);
    launch(T, L_MOM, 10, MCM_BLOCK, k_momentum<MODE_FUSED, 6, false>, cur);
    launch(T, L_MOM, 11, MCM_BLOCK, k_momentum<MODE_FVEL, 4, false>, cur);
    launch(T, L_MOM, 12, MCM_BLOCK, k_momentum<MODE_FVEL, 5, false>, cur);
    launch(T, L_MOM, 13, MCM_BLOCK, k_momentum<MODE_FVEL, 6, false>, cur);
    launch(T, L_XSPH, 0, MCM_BLOCK, k_xsph, cur);
    launch(T, L_XSPH, 0, MCM_BLOCK, k_xsph_move, cur);
}

// the ghosts' stress is refreshed first, the energy and bounds fold follows
void launch_momentum(const Target& T, int cur) {
    launch_ghost_stress(T, cur);
    launch_momentum_kernel(T, cur);
    launch_c_fold(T, 1);
}

void launch_step(const Target& T, int cur) {
    launch_neighbours(T, cur);
    launch_strain(T, cur);
    launch_constitutive(T, cur);
    launch_momentum(T, cur);
}

int get_scalars_raw(mcm_gpu_ctx* ctx, Scalars* out) {
    HIPCHK(hipMemcpyAsync(out, ctx->sc, sizeof(Scalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int put_scalars_raw(mcm_gpu_ctx* ctx, const Scalars& in) {
    HIPCHK(hipMemcpyAsync(ctx->sc, &in, sizeof(Scalars), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// Stages of one loop body of mcm_solution.f90:41-116, one bit each in
// mcm_gpu_ctx::done once run in the current step (cleared by the move).
enum : unsigned {
    ST_STRAIN = 1u, ST_UPDATE_H = 2u, ST_CONST = 4u, ST_DT = 8u, ST_REP = 16u, ST_MOMENTUM = 32u,
    ST_BOUNDCOND = 64u, ST_VELOCITY = 128u
};

// Order check of a stage entry point: every stage reads what the stages
// before it in the loop body wrote, so a call out of that order (or twice in
// one step, or on a run that a batch steps) is refused with MCM_GPU_ERR_ARG
// and a message instead of computing from stale data.
int stage_check(mcm_gpu_ctx* ctx, const char* stage, bool ok, const char* needs) {
    if (ctx->batch) {
        ctx->err = std::string(stage) + ": the run belongs to a batch; step it with mcm_gpu_batch_step or mcm_gpu_step";
        return MCM_GPU_ERR_ARG;
    }
    if (!ok) {
        ctx->err = std::string(stage) + " called out of order: " + needs;
        return MCM_GPU_ERR_ARG;
    }
    return 0;
}

void drop_graph(mcm_gpu_ctx* ctx) {
    if (ctx->graph2) { (void)hipGraphExecDestroy(ctx->graph2); ctx->graph2 = nullptr; }
}

int build_graph(mcm_gpu_ctx* ctx) {
    drop_graph(ctx);
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    launch_step(one(ctx), 0);
    launch_step(one(ctx), 1);
    HIPCHK(hipStreamEndCapture(ctx->stream, &g));
    HIPCHK(hipGraphInstantiate(&ctx->graph2, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    return 0;
}

// Grow the capacity named by the error flags (neighbour slots, ghost slots,
// cells), keep the particle state, clear the flags.
int regrow(mcm_gpu_ctx* ctx, const Scalars& sc_host) {
    DevState& d = ctx->d;
    int flags = sc_host.error_flags;
    if (flags & 32) {
        ctx->err = "penalty contact: a particle has both real and ghost contact partners; the reference's repulsion "
                   "is undefined there (mcm_get_j_cont_info.f90:44-48 reads the real contact list past its end)";
        return MCM_GPU_ERR_ARG;
    }
    if (flags & 128) {
        ctx->err = "cap model (material model 4): the iterated return of mode 10 or 11 did not reach |f| <= 1e-8 in "
                   "100 iterations (the reference stops, mcm_f3dm25.f90:977-983, 1031-1037)";
        return MCM_GPU_ERR_MODEL;
    }
    drop_graph(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (flags & 2) {   // ghosts: unified arrays grow, real part kept
        int ng_new = d.ngcap * 2 + 1024;
        size_t keep = d.np, u = (size_t)d.np + ng_new;
        int rc = 0;
        for (int k = 0; k < 3; k++) {
            rc |= dregrow(ctx, &d.x0[k], u, keep); rc |= dregrow(ctx, &d.x1[k], u, keep); rc |= dregrow(ctx, &d.v[k], u, keep);
        }
        rc |= dregrow(ctx, &d.mass, u, keep);
        for (int k = 0; k < 2; k++) { rc |= dregrow(ctx, &d.hbuf[k], u, keep); rc |= dregrow(ctx, &d.rbuf[k], u, keep); }
        rc |= dregrow(ctx, &d.pmat, u, keep);
        rc |= dregrow(ctx, &d.rs, 2 * RS_SLOTS * u, 2 * RS_SLOTS * keep);
        rc |= dregrow(ctx, &d.rm, 2 * RM_SLOTS * u, 2 * RM_SLOTS * keep);
        if (d.asym) rc |= dregrow(ctx, &d.rml, 4 * u, 4 * keep);
        rc |= dregrow(ctx, &d.gpar, ng_new, 0);
        rc |= dregrow(ctx, &d.gsgn, ng_new, 0);
        if (rc) return rc;
        d.ngcap = ng_new;
    }
    if (flags & (1 | 4)) {
        free_neighbour_arrays(ctx);
        if (flags & 1) d.ncap = std::max(d.ncap + 16, (int)(sc_host.max_count * 1.25) + 16);
        if (flags & 4) d.cellcap = d.cellcap * 2 + 4096;
        int rc = alloc_neighbour_arrays(ctx);
        if (rc) return rc;
    }
    if (flags & 64) {  // skin lists
        dfree(ctx, d.skin_list);
        d.skin_cap = std::max(d.skin_cap + 16, (int)(sc_host.skin_max * 1.25) + 16);
        int rc = dalloc(ctx, &d.skin_list, (size_t)d.skin_cap * d.npad);
        if (rc) return rc;
    }
    Scalars s2 = sc_host;
    s2.abort = 0;
    s2.error_flags = 0;
    s2.skin_valid = 0;   // the step is repeated from its start: build the skin lists again
    return put_scalars_raw(ctx, s2);
}

// ---------------------------------------------------------- initialisation
// Density, energy, pressure and sound speed of particle i at the start
// (mcm_calcrhopc.f90:31-95 with mcm_eoscalc.f90:27-113).  T is double, or a
// dual number when the adjoint differentiates the initial state with
// respect to material constants.
template <class T, class M>
void init_particle(PStateT<T>& s, const M& Mt, const HostDeck& D, long i, double mass, double pcut) {
    using std::sqrt;
    s.mass = mass;
    s.rho0 = Mt.rho;
    s.rho = D.init_rhoe == 1 ? D.pr[i] : 0.0;
    s.e = D.init_rhoe == 1 ? D.pe[i] : 0.0;
    s.pcut = pcut;
    if (Mt.model == 1 || Mt.model == 3 || Mt.model == 4) {
        if (D.init_rhoe != 1) { s.rho = s.rho0; s.e = 0.0; }
        s.p = 0.0;
        // mcm_calcrhopc.f90:44-67: lambda + 2G, model 4's from mcm_sets4
        s.c = sqrt(Mt.strinput[Mt.model == 4 ? 9 : 1] / s.rho);
    } else {
        const auto* ei = Mt.eosinput;
        if (D.init_rhoe != 1) {
            switch (Mt.eos) {
                case 1: s.rho = s.rho0 / ei[8]; s.e = ei[7] * s.mass / s.rho; break;
                case 4: s.rho = s.rho0 / ei[7]; s.e = ei[6] * s.mass / s.rho; break;
                case 13: s.rho = s.rho0 / ei[3]; s.e = ei[2] * s.mass / s.rho; break;
                case 28: s.rho = s.rho0 / ei[3]; s.e = ei[2] * s.mass; break;
                case 41: s.rho = s.rho0 / ei[5]; s.e = ei[2] * ei[3] * s.mass / s.rho; break;
            }
        }
        eos_p(s, Mt);
        sueos(s, Mt);
        if (Mt.model == 9 && s.p < Mt.strinput[0]) s.p = Mt.strinput[0];
    }
}

// Per-particle state as rows of MCM_STATE_WORDS values, the fields of the
// reference's particle record (mcm_database.f90:159-216) in the order of the
// state dumps: x(3), v(3), a(3), rho, rhoold, h, hold, p, e, etry, einc, c,
// vabs, tracerod, pcut, p_cut, efps, fail, mindist, critts, mass, rho0,
// sigma(9), s(9), alfa(9), q(9), rod(9), spin(9) (3x3 column-major), h0,
// xzero(3).  ghost (when given) receives the ghosts' x, v, mass, h (8 per
// ghost) followed by their rho.
int gather_rows(mcm_gpu_ctx* ctx, const Scalars& s, std::vector<double>& rows, std::vector<double>* ghost) {
    DevState& d = ctx->d;
    const size_t np = d.np;
    const int hb = ctx->phase == 1 ? ctx->cur ^ 1 : ctx->cur;
    hipError_t err = hipSuccess;
    auto get = [&](const double* src, size_t n) {
        std::vector<double> v(n);
        if (!err) err = hipMemcpy(v.data(), src, n * sizeof(double), hipMemcpyDeviceToHost);
        return v;
    };
    const size_t ng = ghost ? (size_t)s.ngp : 0;
    std::vector<double> X[3], V[3], Acc[3], SG[6], SD[6], AL[6], ROD[6], SP[3];
    for (int k = 0; k < 3; k++) {
        X[k] = get(ctx->phase == 1 ? d.x1[k] : d.x0[k], np + ng);
        V[k] = get(d.v[k], np + ng);
        Acc[k] = get(d.a[k], np);
        SP[k] = get(d.spin[k], np);
    }
    for (int k = 0; k < 6; k++) {
        SG[k] = get(d.sig[k], np); SD[k] = get(d.sd[k], np); AL[k] = get(d.alfa[k], np); ROD[k] = get(d.rod[k], np);
    }
    auto rho = get(d.rbuf[hb], np + ng), h = get(d.hbuf[hb], np + ng), mass = get(d.mass, np + ng);
    auto rhoold = get(d.rhoold, np), hold = get(d.hold, np), p = get(d.p, np), e = get(d.e, np);
    auto etry = get(d.etry, np), einc = get(d.einc, np), c = get(d.c, np), vabs = get(d.vabs, np);
    auto tr = get(d.tracerod, np), pcut = get(d.pcut, np), p_cut = get(d.p_cut, np), efps = get(d.efps, np);
    auto fail = get(d.fail, np), mind = get(d.mindist, np), crit = get(d.critts_p, np), rho0 = get(d.rho0, np);
    auto h0 = get(d.h0, np), q = get(d.q, np);
    std::vector<double> SL[3];   // the stress's lower triangle of a run with a cap-model material
    if (d.asym)
        for (int k = 0; k < 3; k++) SL[k] = get(d.sigl[k], np);
    if (err) { ctx->err = std::string("state download: ") + hipGetErrorString(err); return MCM_GPU_ERR_HIP; }
    static const int map[9] = {S11, S12, S13, S12, S22, S23, S13, S23, S33};
    rows.assign(np * MCM_STATE_WORDS, 0.0);
    for (size_t i = 0; i < np; i++) {
        double* row = &rows[i * MCM_STATE_WORDS];
        for (int k = 0; k < 3; k++) { row[k] = X[k][i]; row[3 + k] = V[k][i]; row[6 + k] = Acc[k][i]; }
        row[9] = rho[i]; row[10] = rhoold[i]; row[11] = h[i]; row[12] = hold[i]; row[13] = p[i]; row[14] = e[i];
        row[15] = etry[i]; row[16] = einc[i]; row[17] = c[i]; row[18] = vabs[i]; row[19] = tr[i]; row[20] = pcut[i];
        row[21] = p_cut[i]; row[22] = efps[i]; row[23] = fail[i]; row[24] = mind[i]; row[25] = crit[i];
        row[26] = mass[i]; row[27] = rho0[i];
        for (int k = 0; k < 9; k++) {
            int m = map[k];
            row[28 + k] = SG[m][i]; row[37 + k] = SD[m][i]; row[46 + k] = AL[m][i];
            row[55 + k] = (k == 0 || k == 4 || k == 8) ? q[i] : 0.0;
            row[64 + k] = ROD[m][i];
        }
        if (d.asym) { row[28 + 1] = SL[0][i]; row[28 + 2] = SL[1][i]; row[28 + 5] = SL[2][i]; }
        // spin: column-major (r,c) -> W12 at (0,1)=k3, W13 at (0,2)=k6, W23 at (1,2)=k7
        double w12 = SP[0][i], w13 = SP[1][i], w23 = SP[2][i];
        double sp[9] = {0.0, -w12, -w13, w12, 0.0, -w23, w13, w23, 0.0};
        for (int k = 0; k < 9; k++) row[73 + k] = sp[k];
        row[82] = h0[i];
        for (int k = 0; k < 3; k++) row[83 + k] = ctx->xzero[3 * i + k];
    }
    if (ghost) {
        ghost->assign(9 * ng, 0.0);
        for (size_t g = 0; g < ng; g++) {
            size_t u = np + g;
            double r8[8] = {X[0][u], X[1][u], X[2][u], V[0][u], V[1][u], V[2][u], mass[u], h[u]};
            for (int k = 0; k < 8; k++) (*ghost)[8 * g + k] = r8[k];
            (*ghost)[8 * ng + g] = rho[u];
        }
    }
    return 0;
}

// The inverse of gather_rows for the real particles at a step boundary (h
// and rho into both parity buffers), plus the scalars the next step starts
// from.  The strain records are laid down by the next search.
int scatter_rows(mcm_gpu_ctx* ctx, const double* rows, const mcm_gpu_scalars& sc) {
    DevState& d = ctx->d;
    const size_t np = d.np;
    hipStream_t st = ctx->stream;
    hipError_t err = hipSuccess;
    std::vector<std::vector<double>> keep;   // host columns alive until the copies are done
    auto put = [&](double* dst, int word) {
        if (err) return;
        keep.emplace_back(np);
        std::vector<double>& c = keep.back();
        for (size_t i = 0; i < np; i++) c[i] = rows[i * MCM_STATE_WORDS + word];
        err = hipMemcpyAsync(dst, c.data(), np * sizeof(double), hipMemcpyHostToDevice, st);
    };
    for (int k = 0; k < 3; k++) { put(d.x0[k], k); put(d.v[k], 3 + k); put(d.a[k], 6 + k); }
    for (int b = 0; b < 2; b++) { put(d.rbuf[b], 9); put(d.hbuf[b], 11); }
    put(d.rhoold, 10); put(d.hold, 12); put(d.p, 13); put(d.e, 14); put(d.etry, 15); put(d.einc, 16);
    put(d.c, 17); put(d.vabs, 18); put(d.tracerod, 19); put(d.pcut, 20); put(d.p_cut, 21); put(d.efps, 22);
    put(d.fail, 23); put(d.mindist, 24); put(d.critts_p, 25); put(d.mass, 26); put(d.rho0, 27);
    // symmetric tensors from their upper triangle (column-major (r,c) at r + 3c)
    static const int upper[6] = {0, 4, 8, 3, 6, 7};   // 11 22 33 12 13 23
    for (int k = 0; k < 6; k++) {
        put(d.sig[k], 28 + upper[k]); put(d.sd[k], 37 + upper[k]); put(d.alfa[k], 46 + upper[k]);
        put(d.rod[k], 64 + upper[k]);
    }
    if (d.asym) { put(d.sigl[0], 28 + 1); put(d.sigl[1], 28 + 2); put(d.sigl[2], 28 + 5); }   // 21 31 32
    put(d.q, 55);                                          // q is isotropic: q(1,1)
    put(d.spin[0], 73 + 3); put(d.spin[1], 73 + 6); put(d.spin[2], 73 + 7);   // W12 W13 W23
    put(d.h0, 82);
    // material ids, displacement codes and the material table come from the deck
    const HostDeck& D = ctx->deck;
    if (!err) err = hipMemcpyAsync(d.pmat, D.pmat.data(), np * sizeof(int), hipMemcpyHostToDevice, st);
    if (!err) err = hipMemcpyAsync(d.dispbc, D.dispbc.data(), np * sizeof(int), hipMemcpyHostToDevice, st);
    if (!err) err = hipMemcpyAsync(d.mat, D.mat.data(), D.nummat * sizeof(MatParams), hipMemcpyHostToDevice, st);
    if (!err) err = hipMemsetAsync(d.defps_prev, 0, np * sizeof(double), st);
    if (!err) err = hipMemsetAsync(d.yield, 0, np * sizeof(int), st);
    if (!err) err = hipStreamSynchronize(st);
    if (err) { ctx->err = std::string("state upload: ") + hipGetErrorString(err); return MCM_GPU_ERR_HIP; }
    ctx->xzero.resize(3 * np);
    for (size_t i = 0; i < np; i++)
        for (int k = 0; k < 3; k++) ctx->xzero[3 * i + k] = rows[i * MCM_STATE_WORDS + 83 + k];
    Scalars s{};
    int rc = get_scalars_raw(ctx, &s);
    if (rc) return rc;
    s.dt = sc.dt; s.dtold = sc.dtold; s.ptime = sc.ptime; s.critts = sc.critts; s.hmax = sc.hmax;
    for (int k = 0; k < 3; k++) { s.coord_min[k] = sc.coord_min[k]; s.coord_max[k] = sc.coord_max[k]; }
    s.timestep = sc.timestep;
    s.abort = 0;
    s.error_flags = 0;
    s.skin_valid = 0;   // positions replaced: the skin lists are built again
    if ((rc = put_scalars_raw(ctx, s))) return rc;
    ctx->phase = 0;
    ctx->done = 0;
    ctx->fresh = false;
    return 0;
}

// The neighbour search of mcm_neighbours.f90:29-49 (setuplist, ghost_setup,
// ll_neighbours); a capacity that overflows grows and the search is repeated
// (mcm_ll_neighbours.f90:44-311 stops instead).  Lays down this step's strain
// records.
int search(mcm_gpu_ctx* ctx) {
    for (int attempt = 0; attempt < 8; attempt++) {
        launch_neighbours(one(ctx), ctx->cur);
        HIPCHK(hipGetLastError());
        Scalars s;
        int rc = get_scalars_raw(ctx, &s);
        if (rc) return rc;
        if (!s.abort) {
            ctx->fresh = true;
            return MCM_GPU_OK;
        }
        rc = regrow(ctx, s);
        if (rc) return rc;
    }
    ctx->err = "neighbour search did not fit after regrowing";
    return MCM_GPU_ERR_CAPACITY;
}

// mcm_initial.f90:33-101: the per-particle part runs once on the host, the
// first neighbour search, critts and energy totals on the device.
int initial(mcm_gpu_ctx* ctx) {
    HostDeck& D = ctx->deck;
    DevState& d = ctx->d;
    const long n = D.np;
    {   // state that the upload below does not set starts at zero, as in a new run
        double* zero[] = {d.p_cut, d.efps, d.q, d.mindist, d.tracerod, d.einc, d.etry, d.critts_p, d.defps_prev};
        for (double* z : zero) HIPCHK(hipMemsetAsync(z, 0, n * sizeof(double), ctx->stream));
        for (int k = 0; k < 6; k++) {
            HIPCHK(hipMemsetAsync(d.sd[k], 0, n * sizeof(double), ctx->stream));
            HIPCHK(hipMemsetAsync(d.alfa[k], 0, n * sizeof(double), ctx->stream));
            HIPCHK(hipMemsetAsync(d.rod[k], 0, n * sizeof(double), ctx->stream));
        }
        for (int k = 0; k < 3; k++) {
            HIPCHK(hipMemsetAsync(d.spin[k], 0, n * sizeof(double), ctx->stream));
            HIPCHK(hipMemsetAsync(d.a[k], 0, n * sizeof(double), ctx->stream));
        }
        HIPCHK(hipMemsetAsync(d.yield, 0, n * sizeof(int), ctx->stream));
    }
    static const int mask[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    std::vector<double> v = D.v;
    for (long i = 0; i < n; i++) {   // mcm_init_boundcond.f90:25-54
        int bc = D.dispbc[i];
        if (bc >= 1 && bc <= 7)
            for (int k = 0; k < 3; k++)
                if (mask[bc][k]) v[3 * i + k] = 0.0;
    }
    std::vector<double> hx[3], hv[3];
    for (int k = 0; k < 3; k++) { hx[k].resize(n); hv[k].resize(n); }
    std::vector<double> mass(n), h(n), rho(n), rho0(n), p(n), e(n), c(n), vabs(n), pcut(n), fail(n, 1.0);
    std::vector<double> sig[6];
    for (int k = 0; k < 6; k++) sig[k].assign(n, 0.0);
    Scalars sc{};
    for (int k = 0; k < 3; k++) { sc.coord_min[k] = 1.0e20; sc.coord_max[k] = -1.0e20; }
    ctx->xzero.assign(3 * n, 0.0);
    for (long i = 0; i < n; i++) {
        const MatParams& M = D.mat[D.pmat[i]];
        for (int k = 0; k < 3; k++) {
            hx[k][i] = D.x[3 * i + k];
            hv[k][i] = v[3 * i + k];
            ctx->xzero[3 * i + k] = D.x[3 * i + k];
        }
        mass[i] = D.pm[i];
        // mcm_initvbls.f90:48-91
        if (M.model == 1 || M.model == 3 || M.model == 4) pcut[i] = -9.90E+20;
        else if (M.model == 9) pcut[i] = M.strinput[0];
        else if (M.model == 10) pcut[i] = M.strinput[3];
        double va = 0.0;
        for (int k = 0; k < 3; k++) va = va + hv[k][i] * hv[k][i];
        vabs[i] = std::sqrt(va);
        for (int k = 0; k < 3; k++) {
            if (hx[k][i] < sc.coord_min[k]) sc.coord_min[k] = hx[k][i];
            if (hx[k][i] > sc.coord_max[k]) sc.coord_max[k] = hx[k][i];
        }
        PState s{};
        init_particle(s, M, D, i, mass[i], pcut[i]);
        rho[i] = s.rho; rho0[i] = s.rho0; p[i] = s.p; e[i] = s.e; c[i] = s.c;
        for (int k = 0; k < 3; k++) sig[k][i] = -s.p;   // mcm_initsigma.f90:35-43
        h[i] = D.init_h_opt == 1 ? D.ph[i] : M.h;         // mcm_init_h.f90:22-48
    }
    double hmax = 0.0;
    if (D.init_h_opt == 0) {
        for (int m = 0; m < D.nummat; m++) hmax = std::max(hmax, D.mat[D.pmat[m]].h);
    } else {
        for (long i = 0; i < n; i++) hmax = std::max(hmax, h[i]);
    }
    sc.hmax = hmax;
    sc.timestep = 1;
    hipStream_t st = ctx->stream;
    auto up = [&](double* dst, const std::vector<double>& src) {
        return hipMemcpyAsync(dst, src.data(), n * sizeof(double), hipMemcpyHostToDevice, st);
    };
    hipError_t e2 = hipSuccess;
    for (int k = 0; k < 3; k++) {
        e2 = e2 ? e2 : up(d.x0[k], hx[k]);
        e2 = e2 ? e2 : up(d.v[k], hv[k]);
    }
    e2 = e2 ? e2 : up(d.mass, mass);
    for (int b = 0; b < 2; b++) { e2 = e2 ? e2 : up(d.hbuf[b], h); e2 = e2 ? e2 : up(d.rbuf[b], rho); }
    e2 = e2 ? e2 : up(d.hold, h);
    e2 = e2 ? e2 : up(d.h0, h);
    e2 = e2 ? e2 : up(d.rho0, rho0);
    e2 = e2 ? e2 : up(d.rhoold, rho);
    e2 = e2 ? e2 : up(d.p, p);
    e2 = e2 ? e2 : up(d.e, e);
    e2 = e2 ? e2 : up(d.c, c);
    e2 = e2 ? e2 : up(d.vabs, vabs);
    e2 = e2 ? e2 : up(d.pcut, pcut);
    e2 = e2 ? e2 : up(d.fail, fail);
    for (int k = 0; k < 6; k++) e2 = e2 ? e2 : up(d.sig[k], sig[k]);
    if (!e2) e2 = hipMemcpyAsync(d.pmat, D.pmat.data(), n * sizeof(int), hipMemcpyHostToDevice, st);
    if (!e2) e2 = hipMemcpyAsync(d.dispbc, D.dispbc.data(), n * sizeof(int), hipMemcpyHostToDevice, st);
    if (!e2) e2 = hipMemcpyAsync(d.mat, D.mat.data(), D.nummat * sizeof(MatParams), hipMemcpyHostToDevice, st);
    if (e2) { ctx->err = std::string("upload: ") + hipGetErrorString(e2); return MCM_GPU_ERR_HIP; }
    int rc = put_scalars_raw(ctx, sc);
    if (rc) return rc;
    ctx->cur = 0;
    ctx->phase = 0;
    ctx->done = 0;
    // mcm_init_neighb -> calccritts -> dt -> calctotale
    rc = search(ctx);
    if (rc) return rc;
    k_calccritts<<<ctx->nb, MCM_BLOCK, 0, st>>>(d, 0);
    k_calccritts_fold<<<1, MCM_BLOCK, 0, st>>>(d, ctx->nb);
    Scalars s3;
    rc = get_scalars_raw(ctx, &s3);
    if (rc) return rc;
    s3.timestep = 1;
    if (D.itss != 0.0) s3.dt = std::min(D.itss, s3.critts * D.tssf);
    else s3.dt = s3.critts * D.tssf;
    rc = put_scalars_raw(ctx, s3);
    if (rc) return rc;
    k_rs_shift<<<ctx->nb, MCM_BLOCK, 0, st>>>(d);
    k_calctotale<<<ctx->nb, MCM_BLOCK, 0, st>>>(d);
    launch_c_fold(one(ctx), 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// A context with its device and stream set (stream: the batch's stream for
// a run of a batch, else a stream of its own).
int new_run(const mcm_gpu_options* opt, hipStream_t stream, mcm_gpu_ctx** out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MCM_GPU_ERR_NODEVICE;
    mcm_gpu_ctx* ctx = new mcm_gpu_ctx();
    *out = ctx;
    ctx->device = opt ? opt->device : 0;
    ctx->corr_one_pass = opt ? opt->corrected_one_pass : 0;
    ctx->pair_mapping = opt ? opt->pair_mapping : 0;
    if (ctx->pair_mapping < 0 || ctx->pair_mapping > 1) { ctx->err = "pair_mapping must be 0 or 1"; return MCM_GPU_ERR_ARG; }
    HIPCHK(hipSetDevice(ctx->device));
    if (stream) {
        ctx->stream = stream;
        ctx->own_stream = false;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    }
    return MCM_GPU_OK;
}

// binary deck bytes held in memory (the layout of Deck.write_bin)
int read_deck_memory(mcm_gpu_ctx* ctx, const void* data, size_t len) {
    FILE* f = fmemopen(const_cast<void*>(data), len, "rb");
    if (!f) { ctx->err = "cannot read the deck buffer"; return MCM_GPU_ERR_IO; }
    int rc = read_binary_deck(ctx, f);
    fclose(f);
    return rc ? rc : finish_deck(ctx);
}

// Device state of a run from its deck: sizes, boundary planes, capacities,
// allocations.
int setup_run(mcm_gpu_ctx* ctx, const mcm_gpu_options* opt) {
    HostDeck& D = ctx->deck;
    DevState& d = ctx->d;
    d.np = D.np;
    d.npad = (D.np + 63) / 64 * 64;
    d.nummat = D.nummat;
    d.disctype = D.disctype;
    d.h_opt = D.h_opt;
    d.tcrit_opt = D.tcrit_opt;
    d.contacttype = D.contacttype > 0 ? D.contacttype : 0;
    d.veloc_opt = D.veloc_opt;
    // mcm_control_v2.f90:223-243, mcm_baseaccel_in.f90:33-35
    d.drelax = D.drelax_scale > 0.1 && D.drelax_scale < 1.0;
    d.drelax_scale = D.drelax_scale;
    d.baseaccel = D.nthp[0] == 1 || D.nthp[1] == 1 || D.nthp[2] == 1;
    for (int k = 0; k < 3; k++) d.base_a[k] = D.nthp[k] == 1 ? D.base_a[k] : 0.0;
    d.tssf = D.tssf;
    d.pi = acos(-1.0);
    d.ngcap = (opt && opt->ghost_capacity > 0) ? opt->ghost_capacity : D.np / 2 + 4096;
    d.ncap = (opt && opt->nbr_capacity > 0) ? opt->nbr_capacity : 128;
    d.asym = 0;
    for (int m = 0; m < D.nummat; m++) d.asym |= D.mat[m].model == 4 ? 1 : 0;
    // boundary planes (mcm_inp_symmetry_v2.f90:83-137, mcm_init_boundary.f90:28-116)
    d.boundary = 0;
    d.nplanes = 0;
    for (int k = 0; k < 6; k++)
        if (D.mirr_opt == 1 && D.sym_flag[k] != 0) d.boundary = 1;
    for (int dim = 0; dim < 3; dim++) {
        for (int side = 0; side < 2; side++) {
            int code = D.mirr_opt == 1 ? D.sym_flag[2 * dim + side] : 0;
            d.bcode[dim][side] = code;
            d.bx[dim][side] = D.sym_x[2 * dim + side];
        }
        d.btype[dim] = 1 + (d.bcode[dim][0] > 0 ? 1 : 0) + (d.bcode[dim][1] > 0 ? 2 : 0);
    }
    for (int dim = 0; dim < 3; dim++)
        for (int side = 0; side < 2; side++) {
            int code = d.bcode[dim][side];
            if (code != 1 && code != 2) continue;
            PlaneDev& P = d.planes[d.nplanes++];
            P.dim = dim; P.side = side; P.code = code; P.pad_ = 0;
            for (int k = 0; k < 3; k++) { P.mult[k] = 1.0; P.add[k] = 0.0; }
            if (code == 1) {
                P.mult[dim] = -1.0;
                P.add[dim] = 2.0 * d.bx[dim][side];
            } else {
                double span = d.bx[dim][1] - d.bx[dim][0];
                P.add[dim] = side == 0 ? -span : span;
            }
        }
    // cell capacity from the initial bounding box, with room for the bar to spread
    {
        double hmax = 0.0;
        for (int m = 0; m < D.nummat; m++) hmax = std::max(hmax, D.mat[m].h);
        if (D.init_h_opt == 1) for (int i = 0; i < D.np; i++) hmax = std::max(hmax, D.ph[i]);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = 0; i < D.np; i++)
            for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], D.x[3 * i + k]); hi[k] = std::max(hi[k], D.x[3 * i + k]); }
        double cells = 1.0;
        for (int k = 0; k < 3; k++) {
            double span = (hi[k] - lo[k]);
            if (d.bcode[k][0] || d.bcode[k][1]) span = std::max(span, std::fabs(d.bx[k][1] - d.bx[k][0]));
            cells *= (span / (2.0 * hmax) + 4.0) * 2.0;
        }
        d.cellcap = (int)std::min(cells + 4096.0, 2.0e9);
    }
    ctx->nb = blocks_for(D.np, MCM_BLOCK);
    d.nblk = ctx->nb;
    d.skin = opt ? opt->nbr_skin_milli / 1000.0 : 0.0;
    if (opt && opt->nbr_skin_milli < 0) { ctx->err = "nbr_skin_milli must not be negative"; return MCM_GPU_ERR_ARG; }
    if (d.skin > 0.0 && (d.contacttype > 0 || ctx->pair_mapping != 0)) {
        ctx->err = "a neighbour skin needs kernel contact (contacttype 0) and pair_mapping 0";
        return MCM_GPU_ERR_ARG;
    }
    int rc = alloc_all(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MCM_GPU_OK;
}

void destroy_run(mcm_gpu_ctx* ctx) {
    // a recording holds device buffers and a pointer to its run: release it first
    while (!ctx->recs.empty()) mcm_gpu_adj_destroy(ctx->recs.back());
    drop_graph(ctx);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : ctx->allocs) (void)hipFree(p);
    if (ctx->stream && ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

// Mean per-stage times of nsteps steps launched one at a time on T from
// parity cur, with HIP events on T's stream: ms[0] neighbour search,
// ms[1] strain kernel, ms[2] constitutive + fold, ms[3] momentum kernel,
// ms[4] whole step.
// ms[0..5]: neighbour search with its lists, strain pass, constitutive
// update, momentum kernel, whole step, the list build alone
int time_stages(const Target& T, int cur, int nsteps, double* ms, std::string& err) {
    hipEvent_t ev[8] = {};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 8 && !e; k++) e = hipEventCreate(&ev[k]);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int it = 0; it < nsteps && !e; it++) {
        e = hipEventRecord(ev[0], T.stream);
        launch_search(T, cur);
        if (!e) e = hipEventRecord(ev[7], T.stream);
        launch_lists(T, cur);
        if (!e) e = hipEventRecord(ev[1], T.stream);
        launch_strain(T, cur);
        if (!e) e = hipEventRecord(ev[2], T.stream);
        launch_constitutive(T, cur);
        if (!e) e = hipEventRecord(ev[3], T.stream);
        launch_ghost_stress(T, cur);
        if (!e) e = hipEventRecord(ev[4], T.stream);
        launch_momentum_kernel(T, cur);
        if (!e) e = hipEventRecord(ev[5], T.stream);
        launch_c_fold(T, 1);
        if (!e) e = hipEventRecord(ev[6], T.stream);
        if (!e) e = hipEventSynchronize(ev[6]);
        const int pairs[6][2] = {{0, 1}, {1, 2}, {2, 3}, {4, 5}, {0, 6}, {7, 1}};
        for (int k = 0; k < 6 && !e; k++) {
            float t = 0.0f;
            e = hipEventElapsedTime(&t, ev[pairs[k][0]], ev[pairs[k][1]]);
            acc[k] += t;
        }
        cur ^= 1;
    }
    for (int k = 0; k < 6; k++) ms[k] = acc[k] / nsteps;
    for (int k = 0; k < 8; k++)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    if (!e) e = hipGetLastError();
    if (e) { err = std::string("timed steps: ") + hipGetErrorString(e); return MCM_GPU_ERR_HIP; }
    return 0;
}

#define HIPCHK_B(call)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            b->err = std::string(#call) + ": " + hipGetErrorString(e_);                           \
            return MCM_GPU_ERR_HIP;                                                               \
        }                                                                                         \
    } while (0)

void drop_batch_graph(mcm_gpu_batch* b) {
    if (b->graph2) { (void)hipGraphExecDestroy(b->graph2); b->graph2 = nullptr; }
}

// Upload every run's DevState and the block maps of all launch kinds
// (after runs were added or a run's capacities grew).
int batch_upload(mcm_gpu_batch* b) {
    drop_batch_graph(b);
    HIPCHK_B(hipStreamSynchronize(b->stream));
    const int nr = (int)b->inst.size();
    std::vector<DevState> hs(nr);
    std::vector<BlockEntry> ent;
    b->max_planes = 0;
    for (int i = 0; i < nr; i++) {
        hs[i] = b->inst[i]->d;
        b->max_planes = std::max(b->max_planes, hs[i].nplanes);
    }
    for (int kind = 0; kind < L_NKIND; kind++)
        for (int v = 0; v < kNVariant; v++) {
            const size_t off = ent.size();
            for (int i = 0; i < nr; i++) {
                const int n = kind_blocks(b->inst[i], kind, v);
                for (int j = 0; j < n; j++) ent.push_back(BlockEntry{i, j, n, 0});
            }
            if (ent.size() > (size_t)INT32_MAX) { b->err = "batch too large for one launch"; return MCM_GPU_ERR_CAPACITY; }
            b->map_off[kind][v] = (int)off;
            b->map_n[kind][v] = (int)(ent.size() - off);
        }
    if (b->states) { (void)hipFree(b->states); b->states = nullptr; }
    HIPCHK_B(hipMalloc(&b->states, sizeof(DevState) * nr));
    if (ent.size() > b->entries_cap) {
        if (b->entries) (void)hipFree(b->entries);
        b->entries = nullptr;
        b->entries_cap = ent.size() + ent.size() / 4 + 1024;
        HIPCHK_B(hipMalloc(&b->entries, sizeof(BlockEntry) * b->entries_cap));
    }
    HIPCHK_B(hipMemcpyAsync(b->states, hs.data(), sizeof(DevState) * nr, hipMemcpyHostToDevice, b->stream));
    if (!ent.empty())
        HIPCHK_B(hipMemcpyAsync(b->entries, ent.data(), sizeof(BlockEntry) * ent.size(), hipMemcpyHostToDevice, b->stream));
    HIPCHK_B(hipStreamSynchronize(b->stream));
    b->stale = false;
    return 0;
}

int batch_build_graph(mcm_gpu_batch* b) {
    drop_batch_graph(b);
    hipGraph_t g;
    HIPCHK_B(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    launch_step(all(b), 0);
    launch_step(all(b), 1);
    HIPCHK_B(hipStreamEndCapture(b->stream, &g));
    HIPCHK_B(hipGraphInstantiate(&b->graph2, g, nullptr, nullptr, 0));
    HIPCHK_B(hipGraphDestroy(g));
    return 0;
}

// scalar blocks of every run (one synchronisation)
int batch_scalars(mcm_gpu_batch* b, std::vector<Scalars>& out) {
    out.resize(b->inst.size());
    for (size_t i = 0; i < b->inst.size(); i++)
        HIPCHK_B(hipMemcpyAsync(&out[i], b->inst[i]->sc, sizeof(Scalars), hipMemcpyDeviceToHost, b->stream));
    HIPCHK_B(hipStreamSynchronize(b->stream));
    return 0;
}

int check_sticky(mcm_gpu_ctx* ctx) {
    Scalars s;
    int rc = get_scalars_raw(ctx, &s);
    if (rc) return rc;
    if (s.abort) {
        ctx->err = "capacity exceeded (flags " + std::to_string(s.error_flags) + ")";
        return MCM_GPU_ERR_CAPACITY;
    }
    return 0;
}

}  // namespace

// =============================================================== C ABI
extern "C" {

int mcm_gpu_convert_deck(const char* in_path, const char* out_path, char* msg, int32_t msg_len) {
    if (!in_path || !out_path) return MCM_GPU_ERR_ARG;
    mcm_gpu_ctx tmp;
    int rc = read_deck_raw(&tmp, in_path);
    if (rc == 0) {
        rc = write_binary_deck(tmp.deck, out_path);
        if (rc) tmp.err = std::string("cannot write ") + out_path;
    }
    if (msg && msg_len > 0) {
        strncpy(msg, tmp.err.c_str(), (size_t)msg_len - 1);
        msg[msg_len - 1] = 0;
    }
    return rc;
}

int mcm_gpu_create_from_deck(const char* deck_path, const mcm_gpu_options* opt, mcm_gpu_ctx** out) {
    if (!out || !deck_path) return MCM_GPU_ERR_ARG;
    *out = nullptr;
    mcm_gpu_ctx* ctx = nullptr;
    int rc = new_run(opt, nullptr, &ctx);
    *out = ctx;
    if (rc) return rc;
    rc = read_deck(ctx, deck_path);
    return rc ? rc : setup_run(ctx, opt);
}

int mcm_gpu_create_from_memory(const void* data, int64_t len, const mcm_gpu_options* opt, mcm_gpu_ctx** out) {
    if (!out || !data || len <= 0) return MCM_GPU_ERR_ARG;
    *out = nullptr;
    mcm_gpu_ctx* ctx = nullptr;
    int rc = new_run(opt, nullptr, &ctx);
    *out = ctx;
    if (rc) return rc;
    rc = read_deck_memory(ctx, data, (size_t)len);
    return rc ? rc : setup_run(ctx, opt);
}

void mcm_gpu_destroy(mcm_gpu_ctx* ctx) {
    if (!ctx || ctx->batch) return;   // a run of a batch goes with the batch
    destroy_run(ctx);
}

const char* mcm_gpu_last_error(const mcm_gpu_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mcm_gpu_initial(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    return initial(ctx);
}

int mcm_gpu_neighbours(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_neighbours", ctx->phase == 0 && !(ctx->done & ST_STRAIN),
                         "the search starts a step (after mcm_gpu_initial or mcm_gpu_move)");
    return rc ? rc : search(ctx);
}

int mcm_gpu_strain(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_strain", ctx->phase == 0 && ctx->fresh,
                         "the strain pass needs this step's neighbour search (mcm_gpu_neighbours after a move)");
    if (rc) return rc;
    DevState& d = ctx->d;
    switch (corr_variant(ctx)) {
        case 1: k_strain_const<MODE_STRAIN, 1, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur); break;
        case 2: k_strain_const<MODE_STRAIN, 2, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur); break;
        default: k_strain_const<MODE_STRAIN, 0, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur); break;
    }
    HIPCHK(hipGetLastError());
    ctx->phase = 1;
    ctx->done |= ST_STRAIN;
    return 0;
}

int mcm_gpu_update_h(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_update_h", ctx->phase == 1 && !(ctx->done & (ST_UPDATE_H | ST_CONST)),
                         "update_h runs once, after the strain pass and before the constitutive update");
    if (rc) return rc;
    if (ctx->d.h_opt > 0) {
        k_update_h<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, ctx->cur);
        k_update_h_fold<<<1, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, ctx->nb);
    }
    HIPCHK(hipGetLastError());
    ctx->done |= ST_UPDATE_H;
    return 0;
}

int mcm_gpu_constitutive(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_constitutive", ctx->phase == 1 && !(ctx->done & ST_CONST),
                         "the constitutive update runs once, after the strain pass");
    if (rc) return rc;
    k_strain_const<MODE_CONST, 0, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, ctx->cur);
    k_strain_fold<MODE_CONST><<<1, MCM_BLOCK, 0, ctx->stream>>>(ctx->d);
    HIPCHK(hipGetLastError());
    ctx->done |= ST_CONST;
    return 0;
}

int mcm_gpu_rep_cont(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_rep_cont",
                         ctx->phase == 1 && (ctx->done & ST_DT) && !(ctx->done & (ST_REP | ST_MOMENTUM)),
                         "the contact repulsion runs once, after the dt update and before the momentum stage");
    if (rc) return rc;
    if (ctx->d.contacttype > 0) k_rep_cont<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, ctx->cur);
    HIPCHK(hipGetLastError());
    ctx->done |= ST_REP;
    return 0;
}

// dt update: after the constitutive update (the reference's place), or at a
// step start before the strain pass
static int time_check(mcm_gpu_ctx* ctx, const char* stage) {
    return stage_check(ctx, stage, !(ctx->done & ST_DT) && (ctx->phase == 0 || (ctx->done & ST_CONST)),
                       "the dt update runs once per step, after the constitutive update or before the strain pass");
}

// after a new dt at a step start, the strain records of this step's search
// take their half step with it (a later search lays them down itself)
static int time_done(mcm_gpu_ctx* ctx) {
    if (ctx->phase == 0 && ctx->fresh) k_rs_shift<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d);
    HIPCHK(hipGetLastError());
    ctx->done |= ST_DT;
    return 0;
}

int mcm_gpu_dt_update(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = time_check(ctx, "mcm_gpu_dt_update");
    if (rc) return rc;
    k_dt_update<<<1, 1, 0, ctx->stream>>>(ctx->d);
    return time_done(ctx);
}

int mcm_gpu_set_time(mcm_gpu_ctx* ctx, double dt, double dtold, double ptime) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    if (!(dt >= 0.0) || !(dtold >= 0.0) || !std::isfinite(dt) || !std::isfinite(dtold) || !std::isfinite(ptime)) {
        ctx->err = "mcm_gpu_set_time: dt and dtold must be finite and not negative, ptime finite";
        return MCM_GPU_ERR_ARG;
    }
    // absolute values, so a repeated call changes nothing; after the momentum
    // stage has used dt it would split the step between two values
    int rc = stage_check(ctx, "mcm_gpu_set_time",
                         ctx->phase == 0 || ((ctx->done & ST_CONST) && !(ctx->done & ST_MOMENTUM)),
                         "the host's dt is taken at a step start or between the constitutive update and the "
                         "momentum stage");
    if (rc) return rc;
    Scalars s;
    if ((rc = get_scalars_raw(ctx, &s))) return rc;
    s.dt = dt;
    s.dtold = dtold;
    s.ptime = ptime;
    if ((rc = put_scalars_raw(ctx, s))) return rc;
    return time_done(ctx);
}

int mcm_gpu_momentum(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_momentum",
                         ctx->phase == 1 && (ctx->done & ST_CONST) && (ctx->done & ST_DT) && !(ctx->done & ST_MOMENTUM),
                         "the momentum stage runs once, after the constitutive update and the dt update");
    if (rc) return rc;
    DevState& d = ctx->d;
    if (d.boundary) k_ghost_stress<<<ghost_blocks(d), MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur);
    switch (corr_variant(ctx) + (d.asym ? 4 : 0)) {
        case 1: k_momentum<MODE_MOMENTUM, 1, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur); break;
        case 2: k_momentum<MODE_MOMENTUM, 2, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur); break;
        case 4: k_momentum<MODE_MOMENTUM, 4, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur); break;
        case 5: k_momentum<MODE_MOMENTUM, 5, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur); break;
        case 6: k_momentum<MODE_MOMENTUM, 6, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur); break;
        default: k_momentum<MODE_MOMENTUM, 0, true><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(d, ctx->cur); break;
    }
    HIPCHK(hipGetLastError());
    ctx->done |= ST_MOMENTUM;
    return 0;
}

int mcm_gpu_boundcond(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_boundcond",
                         (ctx->done & ST_MOMENTUM) && !(ctx->done & (ST_BOUNDCOND | ST_VELOCITY)),
                         "the boundary conditions apply once, to the momentum stage's acceleration");
    if (rc) return rc;
    k_boundcond<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d);
    HIPCHK(hipGetLastError());
    ctx->done |= ST_BOUNDCOND;
    return 0;
}

int mcm_gpu_calctotale(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    // a reduction of the current state: any point of the step
    int rc = stage_check(ctx, "mcm_gpu_calctotale", true, "");
    if (rc) return rc;
    k_calctotale<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d);
    launch_c_fold(one(ctx), 0);
    HIPCHK(hipGetLastError());
    return 0;
}

int mcm_gpu_update_velocity(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_update_velocity", (ctx->done & ST_MOMENTUM) && !(ctx->done & ST_VELOCITY),
                         "the velocity update runs once, after the momentum stage");
    if (rc) return rc;
    k_update_velocity<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d);
    if (ctx->d.veloc_opt == 1) {   // XSPH, mcm_update_velocity.f90:61-87
        k_xsph<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, ctx->cur);
        k_xsph_apply<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d);
    }
    HIPCHK(hipGetLastError());
    ctx->done |= ST_VELOCITY;
    return 0;
}

int mcm_gpu_move(mcm_gpu_ctx* ctx) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_move", ctx->phase == 1 && (ctx->done & ST_VELOCITY),
                         "the move ends a step, after the velocity update");
    if (rc) return rc;
    k_move<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d);
    k_move_fold<<<1, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, ctx->nb);
    HIPCHK(hipGetLastError());
    ctx->cur ^= 1;
    ctx->phase = 0;
    ctx->done = 0;
    ctx->fresh = false;   // positions moved: the next strain pass needs a new search
    return 0;
}

int mcm_gpu_calccritts(mcm_gpu_ctx* ctx, double* critts_out) {
    if (!ctx) return MCM_GPU_ERR_ARG;
    int rc = stage_check(ctx, "mcm_gpu_calccritts", true, "");
    if (rc) return rc;
    int hb = ctx->phase == 1 ? ctx->cur ^ 1 : ctx->cur;
    k_calccritts<<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, hb);
    k_calccritts_fold<<<1, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, ctx->nb);
    HIPCHK(hipGetLastError());
    Scalars s;
    rc = get_scalars_raw(ctx, &s);
    if (rc) return rc;
    if (critts_out) *critts_out = s.critts;
    return 0;
}

int mcm_gpu_step(mcm_gpu_ctx* ctx, int32_t nsteps) {
    if (!ctx || nsteps < 0) return MCM_GPU_ERR_ARG;
    if (ctx->phase != 0 || ctx->done != 0) {
        ctx->err = "mcm_gpu_step: the run is inside a step begun with the stage entry points";
        return MCM_GPU_ERR_ARG;
    }
    // the fused steps search again themselves; a run of a batch may grow its
    // capacities here, so the batch takes its states afresh
    ctx->fresh = false;
    if (ctx->batch) ctx->batch->stale = true;
    Scalars s0;
    int rc = get_scalars_raw(ctx, &s0);
    if (rc) return rc;
    const int t_start = s0.timestep;
    const int cur0 = ctx->cur;
    int done = 0;
    for (int attempt = 0; attempt < 16; attempt++) {
        if (!ctx->graph2) {
            rc = build_graph(ctx);
            if (rc) return rc;
        }
        for (int k = done; k < nsteps;) {
            if (ctx->cur == 0 && nsteps - k >= 2) {
                HIPCHK(hipGraphLaunch(ctx->graph2, ctx->stream));
                k += 2;
            } else {
                launch_step(one(ctx), ctx->cur);
                HIPCHK(hipGetLastError());
                ctx->cur ^= 1;
                k += 1;
            }
        }
        Scalars s;
        rc = get_scalars_raw(ctx, &s);
        if (rc) return rc;
        // the step counter advances only for steps that completed, so a
        // step that hit a capacity limit is redone from its start
        done = s.timestep - t_start;
        ctx->cur = cur0 ^ (done & 1);
        if (!s.abort) return MCM_GPU_OK;
        rc = regrow(ctx, s);
        if (rc) return rc;
    }
    ctx->err = "fused step did not fit after regrowing";
    return MCM_GPU_ERR_CAPACITY;
}

int mcm_gpu_get_scalars(mcm_gpu_ctx* ctx, mcm_gpu_scalars* out) {
    if (!ctx || !out) return MCM_GPU_ERR_ARG;
    Scalars s;
    int rc = get_scalars_raw(ctx, &s);
    if (rc) return rc;
    memset(out, 0, sizeof(*out));
    out->dt = s.dt; out->dtold = s.dtold; out->ptime = s.ptime; out->critts = s.critts; out->hmax = s.hmax;
    out->kinetice = s.kinetice; out->internale = s.internale; out->totale = s.totale;
    for (int k = 0; k < 3; k++) { out->coord_min[k] = s.coord_min[k]; out->coord_max[k] = s.coord_max[k]; }
    out->np = ctx->d.np; out->ngp = s.ngp; out->max_count = s.max_count; out->timestep = s.timestep;
    out->error_flags = s.error_flags;
    out->nummat = ctx->d.nummat;
    out->skin_builds = s.skin_builds;
    return 0;
}

static int copy_d2h(mcm_gpu_ctx* ctx, double* dst, const double* src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return 0;
}

int mcm_gpu_download(mcm_gpu_ctx* ctx, int32_t field, double* host, int64_t n) {
    if (!ctx || !host) return MCM_GPU_ERR_ARG;
    DevState& d = ctx->d;
    const size_t np = d.np;
    const int hb = ctx->phase == 1 ? ctx->cur ^ 1 : ctx->cur;
    std::vector<double> tmp(np);
    auto need = [&](size_t k) { return (size_t)n >= k * np; };
    int rc = 0;
    switch (field) {
        case MCM_F_X: case MCM_F_V: case MCM_F_A: {
            if (!need(3)) return MCM_GPU_ERR_ARG;
            double* const* src = field == MCM_F_X ? (ctx->phase == 1 ? d.x1 : d.x0) : (field == MCM_F_V ? d.v : d.a);
            for (int k = 0; k < 3; k++) {
                rc |= copy_d2h(ctx, tmp.data(), src[k], np);
                HIPCHK(hipStreamSynchronize(ctx->stream));
                for (size_t i = 0; i < np; i++) host[3 * i + k] = tmp[i];
            }
            return rc;
        }
        case MCM_F_SIGMA: {
            if (!need(9)) return MCM_GPU_ERR_ARG;
            static const int map[9] = {S11, S12, S13, S12, S22, S23, S13, S23, S33};
            for (int k = 0; k < 6; k++) {
                rc |= copy_d2h(ctx, tmp.data(), d.sig[k], np);
                HIPCHK(hipStreamSynchronize(ctx->stream));
                for (int e = 0; e < 9; e++)
                    if (map[e] == k)
                        for (size_t i = 0; i < np; i++) host[9 * i + e] = tmp[i];
            }
            return rc;
        }
        default: break;
    }
    if (!need(1)) return MCM_GPU_ERR_ARG;
    const double* src = nullptr;
    switch (field) {
        case MCM_F_RHO: src = d.rbuf[hb]; break;
        case MCM_F_H: src = d.hbuf[hb]; break;
        case MCM_F_P: src = d.p; break;
        case MCM_F_E: src = d.e; break;
        case MCM_F_C: src = d.c; break;
        case MCM_F_EFPS: src = d.efps; break;
        case MCM_F_VABS: src = d.vabs; break;
        case MCM_F_MASS: src = d.mass; break;
        default: return MCM_GPU_ERR_ARG;
    }
    rc = copy_d2h(ctx, host, src, np);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return rc;
}

int mcm_gpu_download_neighbours(mcm_gpu_ctx* ctx, int32_t* counts, int32_t* gcounts, int32_t* ids, int64_t ids_len,
                                int32_t* cap) {
    if (!ctx || !counts || !gcounts || !cap) return MCM_GPU_ERR_ARG;
    DevState& d = ctx->d;
    const size_t np = d.np;
    std::vector<int> nr(np), nt(np), pos(ids ? np : 0), lst(ids ? (size_t)d.ncap * d.npad : 0);
    HIPCHK(hipMemcpyAsync(nr.data(), d.nreal, np * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(nt.data(), d.ntot, np * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (ids) HIPCHK(hipMemcpyAsync(pos.data(), d.spos, np * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (ids) HIPCHK(hipMemcpyAsync(lst.data(), d.nbr, lst.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *cap = d.ncap;
    if (ids && ids_len < (int64_t)np * d.ncap) return MCM_GPU_ERR_ARG;
    for (size_t i = 0; i < np; i++) {
        counts[i] = nr[i];
        gcounts[i] = nt[i] - nr[i];
        if (ids) {
            const size_t t = pos[i];   // lists are stored by sorted position
            for (int k = 0; k < nt[i] && k < d.ncap; k++) ids[i * d.ncap + k] = lst[((t >> 6) * d.ncap + k) * 64 + (t & 63)];
        }
    }
    return 0;
}

int mcm_gpu_get_energies(mcm_gpu_ctx* ctx, double* out, int32_t n) {
    if (!ctx || !out || n < 5 * ctx->d.nummat) return MCM_GPU_ERR_ARG;
    Scalars s;
    int rc = get_scalars_raw(ctx, &s);
    if (rc) return rc;
    for (int m = 0; m < ctx->d.nummat; m++) {
        out[5 * m] = s.mat_ke[m];
        out[5 * m + 1] = s.mat_ie[m];
        for (int k = 0; k < 3; k++) out[5 * m + 2 + k] = s.mat_mom[m][k];
    }
    return 0;
}

int mcm_gpu_dump(mcm_gpu_ctx* ctx, const char* path, int32_t step, int32_t stage) {
    if (!ctx || !path) return MCM_GPU_ERR_ARG;
    DevState& d = ctx->d;
    const size_t np = d.np;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    Scalars s;
    int rc = get_scalars_raw(ctx, &s);
    if (rc) return rc;
    auto geti = [&](const int* src, size_t n) {
        std::vector<int> v(n);
        (void)hipMemcpy(v.data(), src, n * sizeof(int), hipMemcpyDeviceToHost);
        return v;
    };
    const size_t ng = s.ngp;
    std::vector<double> rows, ghost;
    if ((rc = gather_rows(ctx, s, rows, &ghost))) return rc;
    auto nreal = geti(d.nreal, np), ntot = geti(d.ntot, np), cell_of = geti(d.cell_of, np), spos = geti(d.spos, np);
    auto lst = geti(d.nbr, (size_t)d.ncap * d.npad);
    auto gpar = geti(d.gpar, ng), pmat = geti(d.pmat + np, ng);
    FILE* f = fopen(path, "wb");
    if (!f) { ctx->err = "cannot write dump"; return MCM_GPU_ERR_IO; }
    int32_t hdr[6] = {0x4D434D44, step, stage, (int32_t)np, (int32_t)ng, d.nummat};
    fwrite(hdr, 4, 6, f);
    double scv[8] = {s.dt, s.dtold, s.ptime, s.critts, s.hmax, s.kinetice, s.internale, s.totale};
    fwrite(scv, 8, 8, f);
    double gr[12];
    for (int k = 0; k < 3; k++) { gr[2 * k] = s.coord_min[k]; gr[2 * k + 1] = s.coord_max[k]; }
    for (int k = 0; k < 3; k++) { gr[6 + k] = s.grid.gridmin[k]; gr[9 + k] = s.grid.gridsize[k]; }
    fwrite(gr, 8, 12, f);
    fwrite(s.grid.lim, 4, 3, f);
    for (int m = 0; m < d.nummat; m++) {
        double mm[5] = {s.mat_ke[m], s.mat_ie[m], s.mat_mom[m][0], s.mat_mom[m][1], s.mat_mom[m][2]};
        fwrite(mm, 8, 5, f);
    }
    fwrite(rows.data(), 8, rows.size(), f);
    std::vector<int32_t> nn(np), gn(np), llp(np);
    for (size_t i = 0; i < np; i++) { nn[i] = nreal[i]; gn[i] = ntot[i] - nreal[i]; llp[i] = cell_of[i] >= 0 ? 0 : -1; }
    fwrite(nn.data(), 4, np, f);
    fwrite(gn.data(), 4, np, f);
    fwrite(llp.data(), 4, np, f);
    // entry k of particle i (lists are stored by sorted position)
    auto entry = [&](size_t i, int k) {
        const size_t t = spos[i];
        return lst[((t >> 6) * d.ncap + k) * 64 + (t & 63)];
    };
    for (size_t i = 0; i < np; i++)
        for (int k = 0; k < nreal[i] && k < d.ncap; k++) {
            int32_t v = entry(i, k) + 1;
            fwrite(&v, 4, 1, f);
        }
    if (d.boundary) {
        for (size_t i = 0; i < np; i++)
            for (int k = nreal[i]; k < ntot[i] && k < d.ncap; k++) {
                int32_t v = entry(i, k) - (int32_t)np + 1;
                fwrite(&v, 4, 1, f);
            }
        for (size_t g = 0; g < ng; g++) { int32_t v = gpar[g] + 1; fwrite(&v, 4, 1, f); }
        for (size_t g = 0; g < ng; g++) { int32_t v = pmat[g] + 1; fwrite(&v, 4, 1, f); }
        fwrite(ghost.data(), 8, 8 * ng, f);
        for (size_t g = 0; g < ng; g++) fwrite(&ghost[8 * ng + g], 8, 1, f);
        double z = 0.0;
        for (size_t g = 0; g < ng; g++) fwrite(&z, 8, 1, f);
    }
    if (d.asym) {   // cap-model section: 'CAP4', capa(np), ptype(np)
        std::vector<double> capa(np);
        (void)hipMemcpy(capa.data(), d.capa, np * sizeof(double), hipMemcpyDeviceToHost);
        auto ptype = geti(d.ptype, np);
        int32_t tag = 0x43415034;
        fwrite(&tag, 4, 1, f);
        fwrite(capa.data(), 8, np, f);
        fwrite(ptype.data(), 4, np, f);
    }
    fclose(f);
    return 0;
}

int mcm_gpu_download_state(mcm_gpu_ctx* ctx, double* rows, int64_t np, mcm_gpu_scalars* sc) {
    if (!ctx || (!rows && !sc)) return MCM_GPU_ERR_ARG;
    if (rows && np != ctx->d.np) {
        ctx->err = "mcm_gpu_download_state: np must be the run's particle count";
        return MCM_GPU_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    Scalars s;
    int rc = get_scalars_raw(ctx, &s);
    if (rc) return rc;
    if (rows) {
        std::vector<double> r;
        if ((rc = gather_rows(ctx, s, r, nullptr))) return rc;
        memcpy(rows, r.data(), r.size() * sizeof(double));
    }
    if (sc) rc = mcm_gpu_get_scalars(ctx, sc);
    return rc;
}

int mcm_gpu_upload_state(mcm_gpu_ctx* ctx, const double* rows, int64_t np, const mcm_gpu_scalars* sc) {
    if (!ctx || !rows || !sc) return MCM_GPU_ERR_ARG;
    if (np != ctx->d.np) {
        ctx->err = "mcm_gpu_upload_state: np must be the run's particle count";
        return MCM_GPU_ERR_ARG;
    }
    int rc = stage_check(ctx, "mcm_gpu_upload_state", ctx->phase == 0 && ctx->done == 0,
                         "the state is replaced at a step boundary");
    if (rc) return rc;
    if (!(sc->hmax > 0.0) || !(sc->dt >= 0.0) || !(sc->dtold >= 0.0)) {
        ctx->err = "mcm_gpu_upload_state: hmax must be positive and dt, dtold not negative";
        return MCM_GPU_ERR_ARG;
    }
    return scatter_rows(ctx, rows, *sc);
}

int mcm_gpu_cap_state(mcm_gpu_ctx* ctx, double* capa, int32_t* ptype, int64_t np, int32_t upload) {
    if (!ctx || (upload != 0 && upload != 1)) return MCM_GPU_ERR_ARG;
    DevState& d = ctx->d;
    if (np != d.np) {
        ctx->err = "mcm_gpu_cap_state: np must be the run's particle count";
        return MCM_GPU_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (upload) {
        int rc = stage_check(ctx, "mcm_gpu_cap_state", ctx->phase == 0 && ctx->done == 0,
                             "the state is replaced at a step boundary");
        if (rc) return rc;
    }
    if (!d.asym) {   // no cap-model material: zeros out, nothing to take in
        if (!upload) {
            if (capa) std::fill(capa, capa + np, 0.0);
            if (ptype) std::fill(ptype, ptype + np, 0);
        }
        return MCM_GPU_OK;
    }
    const hipMemcpyKind kind = upload ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    if (capa) HIPCHK(upload ? hipMemcpy(d.capa, capa, np * sizeof(double), kind)
                            : hipMemcpy(capa, d.capa, np * sizeof(double), kind));
    if (ptype) HIPCHK(upload ? hipMemcpy(d.ptype, ptype, np * sizeof(int32_t), kind)
                             : hipMemcpy(ptype, d.ptype, np * sizeof(int32_t), kind));
    return MCM_GPU_OK;
}

int mcm_gpu_time_kernels(mcm_gpu_ctx* ctx, int32_t nsteps, double* ms, int32_t n) {
    if (!ctx || !ms || nsteps <= 0 || n < 5 || ctx->phase != 0 || ctx->done != 0) return MCM_GPU_ERR_ARG;
    if (ctx->batch) return stage_check(ctx, "mcm_gpu_time_steps", false, "");
    ctx->fresh = false;
    double t[6];
    int rc = time_stages(one(ctx), ctx->cur, nsteps, t, ctx->err);
    ctx->cur ^= nsteps & 1;
    for (int k = 0; k < std::min(n, 6); k++) ms[k] = t[k];
    return rc ? rc : check_sticky(ctx);
}

int mcm_gpu_time_steps(mcm_gpu_ctx* ctx, int32_t nsteps, double* ms) { return mcm_gpu_time_kernels(ctx, nsteps, ms, 5); }

// ------------------------------------------------------------------ output
static_assert(OUT_WORDS == MCM_OUT_WORDS && OUT_WORDS == vtkout::kWords, "packed output record size");

int mcm_gpu_output_fields(mcm_gpu_ctx* ctx, int32_t precision, void* host, int64_t nbytes) {
    if (!ctx || !host || (precision != 4 && precision != 8)) return MCM_GPU_ERR_ARG;
    const size_t np = ctx->d.np;
    const size_t bytes = np * OUT_WORDS * (size_t)precision;
    if ((size_t)nbytes < bytes) { ctx->err = "output buffer too small"; return MCM_GPU_ERR_ARG; }
    if (!ctx->out_buf) {
        double* p = nullptr;
        int rc = dalloc(ctx, &p, np * OUT_WORDS);
        if (rc) return rc;
        ctx->out_buf = p;
    }
    const int hb = ctx->phase == 1 ? ctx->cur ^ 1 : ctx->cur;
    if (precision == 8) k_pack_output<double><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, hb, ctx->phase, (double*)ctx->out_buf);
    else k_pack_output<float><<<ctx->nb, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, hb, ctx->phase, (float*)ctx->out_buf);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, ctx->out_buf, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MCM_GPU_OK;
}

int mcm_vtk_write(const char* path, int64_t np, const double* fields, const int32_t* material_ids, int32_t material) {
    if (!path || np < 0 || (np > 0 && !fields) || material < 0 || (material > 0 && !material_ids)) return MCM_GPU_ERR_ARG;
    return vtkout::write_vtk(path, (long)np, fields, material_ids, material) == 0 ? MCM_GPU_OK : MCM_GPU_ERR_IO;
}

int mcm_gpu_write_vtk(mcm_gpu_ctx* ctx, const char* path, int32_t material) {
    // a material without particles gives a file without points, as the
    // reference's WriteVTKmat(2) does for a one-material run
    if (!ctx || !path || material < 0) return MCM_GPU_ERR_ARG;
    const size_t np = ctx->d.np;
    std::vector<double> f(np * OUT_WORDS);
    int rc = mcm_gpu_output_fields(ctx, 8, f.data(), (int64_t)(f.size() * sizeof(double)));
    if (rc) return rc;
    std::vector<int32_t> mat(np);
    for (size_t i = 0; i < np; i++) mat[i] = ctx->deck.pmat[i] + 1;
    std::vector<int32_t> ptype;
    if (ctx->d.asym) {
        ptype.resize(np);
        HIPCHK(hipMemcpy(ptype.data(), ctx->d.ptype, np * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    rc = vtkout::write_vtk(path, (long)np, f.data(), mat.data(), material, ptype.empty() ? nullptr : ptype.data()) == 0
             ? MCM_GPU_OK : MCM_GPU_ERR_IO;
    if (rc) ctx->err = std::string("cannot write ") + path;
    return rc;
}

// ------------------------------------------------------------- batched runs
int mcm_gpu_batch_create(const mcm_gpu_options* opt, mcm_gpu_batch** out) {
    if (!out) return MCM_GPU_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MCM_GPU_ERR_NODEVICE;
    mcm_gpu_batch* b = new mcm_gpu_batch();
    *out = b;
    if (opt) b->opt = *opt;
    b->device = b->opt.device;
    HIPCHK_B(hipSetDevice(b->device));
    HIPCHK_B(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    return MCM_GPU_OK;
}

// a run made from a deck file (path) or binary deck bytes (data, len)
static int batch_add(mcm_gpu_batch* b, const char* path, const void* data, int64_t len) {
    if (!b) return MCM_GPU_ERR_ARG;
    mcm_gpu_ctx* ctx = nullptr;
    int rc = new_run(&b->opt, b->stream, &ctx);
    if (!rc) rc = path ? read_deck(ctx, path) : read_deck_memory(ctx, data, (size_t)len);
    if (!rc) rc = setup_run(ctx, &b->opt);
    if (rc) {
        b->err = "run " + std::to_string(b->inst.size()) + ": " + (ctx ? ctx->err : std::string("no context"));
        if (ctx) destroy_run(ctx);
        return rc;
    }
    ctx->batch = b;
    ctx->cur = b->cur;
    b->inst.push_back(ctx);
    b->stale = true;
    return MCM_GPU_OK;
}

int mcm_gpu_batch_add_deck(mcm_gpu_batch* b, const char* deck_path) {
    if (!deck_path) return MCM_GPU_ERR_ARG;
    return batch_add(b, deck_path, nullptr, 0);
}

int mcm_gpu_batch_add_memory(mcm_gpu_batch* b, const void* data, int64_t len) {
    if (!data || len <= 0) return MCM_GPU_ERR_ARG;
    return batch_add(b, nullptr, data, len);
}

int32_t mcm_gpu_batch_size(const mcm_gpu_batch* b) { return b ? (int32_t)b->inst.size() : 0; }

mcm_gpu_ctx* mcm_gpu_batch_run(mcm_gpu_batch* b, int32_t k) {
    if (!b || k < 0 || k >= (int32_t)b->inst.size()) return nullptr;
    return b->inst[k];
}

const char* mcm_gpu_batch_last_error(const mcm_gpu_batch* b) { return b ? b->err.c_str() : "null batch"; }

int mcm_gpu_batch_initial(mcm_gpu_batch* b) {
    if (!b || b->inst.empty()) return MCM_GPU_ERR_ARG;
    for (size_t i = 0; i < b->inst.size(); i++) {
        int rc = initial(b->inst[i]);
        if (rc) { b->err = "run " + std::to_string(i) + ": " + b->inst[i]->err; return rc; }
    }
    b->cur = 0;
    b->stale = true;
    return MCM_GPU_OK;
}

int mcm_gpu_batch_step(mcm_gpu_batch* b, int32_t nsteps) {
    if (!b || nsteps < 0 || b->inst.empty()) return MCM_GPU_ERR_ARG;
    for (size_t i = 0; i < b->inst.size(); i++)
        if (b->inst[i]->phase != 0 || b->inst[i]->cur != b->cur) {
            b->err = "run " + std::to_string(i) + " is not at a step boundary of the batch";
            return MCM_GPU_ERR_ARG;
        }
    std::vector<Scalars> s0, s1;
    int rc = batch_scalars(b, s0);
    if (rc) return rc;
    if (b->stale && (rc = batch_upload(b))) return rc;
    if (!b->graph2 && (rc = batch_build_graph(b))) return rc;
    const int cur0 = b->cur;
    for (int k = 0; k < nsteps;) {
        if (b->cur == 0 && nsteps - k >= 2) {
            HIPCHK_B(hipGraphLaunch(b->graph2, b->stream));
            k += 2;
        } else {
            launch_step(all(b), b->cur);
            HIPCHK_B(hipGetLastError());
            b->cur ^= 1;
            k += 1;
        }
    }
    rc = batch_scalars(b, s1);
    if (rc) return rc;
    // A run that exceeded a capacity stopped at the start of that step (its
    // step counter tells where); it grows and finishes the steps on its own.
    for (size_t i = 0; i < b->inst.size(); i++) {
        mcm_gpu_ctx* r = b->inst[i];
        const int done = s1[i].timestep - s0[i].timestep;
        r->cur = cur0 ^ (done & 1);
        if (done < nsteps) {
            rc = mcm_gpu_step(r, nsteps - done);
            if (rc) { b->err = "run " + std::to_string(i) + ": " + r->err; return rc; }
            b->stale = true;
        }
    }
    return MCM_GPU_OK;
}

int mcm_gpu_batch_time_steps(mcm_gpu_batch* b, int32_t nsteps, double* ms) {
    if (!b || !ms || nsteps <= 0 || b->inst.empty()) return MCM_GPU_ERR_ARG;
    int rc = 0;
    if (b->stale && (rc = batch_upload(b))) return rc;
    double t[6];
    rc = time_stages(all(b), b->cur, nsteps, t, b->err);
    for (int k = 0; k < 5; k++) ms[k] = t[k];
    b->cur ^= nsteps & 1;
    for (mcm_gpu_ctx* r : b->inst) r->cur = b->cur;
    if (rc) return rc;
    std::vector<Scalars> s;
    if ((rc = batch_scalars(b, s))) return rc;
    for (size_t i = 0; i < s.size(); i++)
        if (s[i].abort) {
            b->err = "run " + std::to_string(i) + ": capacity exceeded during timed steps";
            return MCM_GPU_ERR_CAPACITY;
        }
    return MCM_GPU_OK;
}

int mcm_gpu_batch_tune(mcm_gpu_batch* b, int32_t nsteps, int32_t* mapping) {
    if (!b || nsteps <= 0 || b->inst.empty()) return MCM_GPU_ERR_ARG;
    double best = 1.0e300;
    int best_m = 0;
    for (int m = 0; m <= 1; m++) {
        b->opt.pair_mapping = m;
        for (mcm_gpu_ctx* r : b->inst) r->pair_mapping = m;
        b->stale = true;
        int rc = mcm_gpu_batch_step(b, 2);          // block maps, graph, warm caches
        if (rc) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        if ((rc = mcm_gpu_batch_step(b, nsteps))) return rc;   // returns after the stream drained
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (t < best) { best = t; best_m = m; }
    }
    b->opt.pair_mapping = best_m;
    for (mcm_gpu_ctx* r : b->inst) r->pair_mapping = best_m;
    b->stale = true;
    if (mapping) *mapping = best_m;
    return MCM_GPU_OK;
}

void mcm_gpu_batch_destroy(mcm_gpu_batch* b) {
    if (!b) return;
    drop_batch_graph(b);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (mcm_gpu_ctx* r : b->inst) destroy_run(r);
    if (b->states) (void)hipFree(b->states);
    if (b->entries) (void)hipFree(b->entries);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"

#include "mcm_noalloc.hip"
#include "mcm_adjoint.hip"
