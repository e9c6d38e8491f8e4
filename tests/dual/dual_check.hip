// Host check of the dual-number derivatives used by the adjoint: the
// per-particle chain of a model-10 step (rhoupd, stress_rot1, f3dm10, sueos,
// nodalav, hieupd, eqos, stress_up, critts) is run with Dual<4> seeds and
// compared with central differences of the same chain in double, for random
// states (elastic and plastic, compression and tension).  Prints one line
// per case: worst relative error over the outputs and inputs, and the count
// of checked entries; then the count of dual values that differ from the
// double chain's (0: the dual chain takes the forward's branches).
#include <cstdio>
#include <cmath>
#include <random>
#include "../../mcmtop_amd/csrc/mcm_physics.h"

constexpr int NIN = 24;   // rod 6, spin 3, sigma 6, rho, e, efps, |v|, dt, G, sigma_y, Eh, S1
constexpr int NOUT = 11;  // sigma 6, rho, e, efps, q, critts

struct Base {
    double in[NIN];
    double mass, rho0, h, pcut, fail, defps;
    MatParams m;
};

template <class T>
void chain(const Base& b, const T in[NIN], T out[NOUT]) {
    PStateT<T> s;
    MatParamsT<T> m;
    m.model = 10; m.eos = 4; m.visc_type = 1; m.ispall = b.m.ispall; m.efps_flag = 0; m.pad_ = 0;
    m.rho = b.m.rho; m.av_l = b.m.av_l; m.av_q = b.m.av_q; m.h = b.m.h; m.rho_max = b.m.rho_max;
    for (int k = 0; k < 48; k++) m.strinput[k] = b.m.strinput[k];
    for (int k = 0; k < 24; k++) m.eosinput[k] = b.m.eosinput[k];
    m.strinput[0] = in[20]; m.strinput[1] = in[21]; m.strinput[2] = in[22]; m.eosinput[1] = in[23];
    m.g = m.strinput[0];
    s.mass = b.mass; s.rho0 = b.rho0; s.h = b.h; s.hold = b.h; s.h0 = b.h; s.p = 0.0; s.c = 0.0;
    s.pcut = b.pcut; s.p_cut = 0.0; s.fail = b.fail; s.q = 0.0; s.einc = 0.0; s.etry = 0.0;
    for (int c = 0; c < 6; c++) s.rod[c] = in[c];
    for (int c = 0; c < 3; c++) s.spin[c] = in[6 + c];
    for (int c = 0; c < 6; c++) s.sig[c] = in[9 + c];
    s.rho = in[15]; s.e = in[16]; s.efps = in[17]; s.vabs = in[18];
    T dt = in[19];
    s.tracerod = s.rod[S11] + s.rod[S22] + s.rod[S33];
    rhoupd(s, m, dt);
    stress_rot1(s, dt);
    double dprev = b.defps;
    f3dm10(s, m, dt, dprev);
    sueos(s, m);
    nodalav(s, m);
    hieupd(s, dt);
    eqos(s, m, dt);
    stress_up(s);
    T cr = critts(s, m, 1, 0.0);
    for (int c = 0; c < 6; c++) out[c] = s.sig[c];
    out[6] = s.rho; out[7] = s.e; out[8] = s.efps; out[9] = s.q; out[10] = cr;
}

int main() {
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    double worst_all = 0.0;
    long checked_all = 0, value_mismatch = 0;
    const char* names[4] = {"elastic-compression", "plastic-compression", "elastic-tension", "plastic-shear"};
    for (int cs = 0; cs < 4; cs++) {
        double worst = 0.0;
        long checked = 0;
        for (int trial = 0; trial < 25; trial++) {
            Base b{};
            b.m.rho = 8930.0; b.m.av_l = 1.0; b.m.av_q = 2.0; b.m.h = 4.8e-4; b.m.rho_max = 1.0e10;
            b.m.ispall = 1;
            b.m.strinput[3] = -1.0e30;   // pcut default of model 10
            b.m.eosinput[0] = 3940.0; b.m.eosinput[4] = 2.0; b.m.eosinput[5] = 0.47; b.m.eosinput[7] = 1.0;
            b.mass = 8930.0 * 6.4e-11; b.rho0 = 8930.0; b.h = 4.8e-4; b.pcut = -1.0e30; b.fail = 1.0; b.defps = 0.0;
            double* x = b.in;
            const bool plastic = cs == 1 || cs == 3;
            const double rate = plastic ? 4.0e5 : 2.0e3;            // strain rate, 1/s
            for (int c = 0; c < 3; c++) x[c] = (cs == 2 ? 0.5 : -1.0) * rate * (0.6 + 0.4 * U(rng));
            for (int c = 3; c < 6; c++) x[c] = (cs == 3 ? 1.0 : 0.2) * rate * U(rng);
            for (int c = 6; c < 9; c++) x[c] = 0.3 * rate * U(rng);
            for (int c = 9; c < 12; c++) x[c] = 2.0e8 * U(rng) - (cs == 2 ? -1e8 : 3e8);
            for (int c = 12; c < 15; c++) x[c] = 1.0e8 * U(rng);
            x[15] = 8930.0 * (cs == 2 ? 0.999 : 1.02 + 0.01 * U(rng));
            x[16] = b.mass * 2.0e3 * (1.5 + U(rng));
            x[17] = plastic ? 0.02 * (1.0 + U(rng)) : 0.0;
            x[18] = 190.0 * (1.0 + 0.1 * U(rng));
            x[19] = 2.0e-8 * (1.0 + 0.2 * U(rng));
            x[20] = 4.3333e10 * (1.0 + 0.05 * U(rng));
            x[21] = 4.0e8 * (1.0 + 0.1 * U(rng));
            x[22] = 1.0009e8 * (1.0 + 0.1 * U(rng));
            x[23] = 1.49 * (1.0 + 0.05 * U(rng));
            // dual Jacobian, 4 columns per pass; its value part must be the
            // double chain's, bit for bit (same branches)
            double o0[NOUT];
            chain(b, x, o0);
            double J[NOUT][NIN];
            for (int p0 = 0; p0 < NIN; p0 += 4) {
                Dual<4> in[NIN], out[NOUT];
                for (int j = 0; j < NIN; j++) {
                    in[j] = x[j];
                    if (j >= p0 && j < p0 + 4) in[j].d[j - p0] = 1.0;
                }
                chain(b, in, out);
                for (int o = 0; o < NOUT; o++) {
                    if (out[o].v != o0[o]) value_mismatch++;
                    for (int k = 0; k < 4 && p0 + k < NIN; k++) J[o][p0 + k] = out[o].d[k];
                }
            }
            // Richardson-extrapolated central differences, relative steps 2e-5
            // and 1e-5 (fourth order); entries compared where both are well
            // above the differences' rounding level
            for (int j = 0; j < NIN; j++) {
                double fdh[2][NOUT];
                for (int r = 0; r < 2; r++) {
                    double xp[NIN], xm[NIN], op[NOUT], om[NOUT];
                    for (int k = 0; k < NIN; k++) { xp[k] = x[k]; xm[k] = x[k]; }
                    const double h = (r == 0 ? 2.0e-5 : 1.0e-5) * (std::fabs(x[j]) > 0.0 ? std::fabs(x[j]) : 1.0);
                    xp[j] += h;
                    xm[j] -= h;
                    chain(b, xp, op);
                    chain(b, xm, om);
                    for (int o = 0; o < NOUT; o++) fdh[r][o] = (op[o] - om[o]) / (xp[j] - xm[j]);
                }
                for (int o = 0; o < NOUT; o++) {
                    const double fd = (4.0 * fdh[1][o] - fdh[0][o]) / 3.0;
                    const double scale = std::fabs(o0[o]) / (std::fabs(x[j]) > 0.0 ? std::fabs(x[j]) : 1.0);
                    if (std::fabs(fd) < 1e-7 * scale && std::fabs(J[o][j]) < 1e-7 * scale) continue;
                    const double err = std::fabs(J[o][j] - fd) / std::max(std::fabs(fd), 1e-3 * scale);
                    worst = std::max(worst, err);
                    checked++;
                }
            }
        }
        printf("%s worst %.3e checked %ld\n", names[cs], worst, checked);
        worst_all = std::max(worst_all, worst);
        checked_all += checked;
    }
    printf("all worst %.3e checked %ld value_mismatch %ld\n", worst_all, checked_all, value_mismatch);
    return 0;
}
