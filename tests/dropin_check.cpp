// The drop-in shim's copies (include/gizmo_dropin.h) on particle types with the member names and
// shapes of GIZMO's particle_data / cell_data (Vec3-like members indexed [k], a Mat3-like velocity
// gradient indexed [k][d], a symmetric tensor with data[6]; some members single precision, as
// MyFloat can be).  Checks, for n particles of seeded values:
//   * pack: every flat array entry equals the member it comes from, at its documented index;
//   * unpack of one stage writes that stage's members and nothing else, with the flat values;
//   * unpack of the force stage reports a wakeup flag exactly when one is set;
//   * the active flags: gas particles (Type 0) of the list only, indices in range only.
// Exit status 0 and "dropin ok" on success; the first mismatch is printed otherwise.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gizmo_dropin.h"

namespace {

template <class T>
struct V3 {
    T x[3];
    T &operator[](int k) { return x[k]; }
    const T &operator[](int k) const { return x[k]; }
};
template <class T>
struct M3 {
    V3<T> row[3];
    V3<T> &operator[](int k) { return row[k]; }
    const V3<T> &operator[](int k) const { return row[k]; }
};
struct Sym {
    double data[6];
};

struct Part {
    short Type;
    short TimeBin;
    V3<double> Pos;
    double Mass;
    float KernelRadius, NumNgb, DrkernNgbFactor, Particle_DivVel;
    int wakeup;
};

struct Cell {
    double Density, Pressure, InternalEnergyPred, DtInternalEnergy;
    V3<double> VelPred, HydroAccel;
    struct {
        V3<double> Density, Pressure;
        M3<double> Velocity;
    } Gradients;
    Sym NV_T;
    V3<double> NV_T_face_weights;
    double ConditionNumber, FaceClosureError;
    float MaxSignalVel;
};

unsigned long long state = 88172645463325252ull;
double rnd()
{
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return (double)(state >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

int fails = 0;
void expect(bool ok, const char *what, long i)
{
    if (!ok && fails++ < 10) std::printf("mismatch: %s at particle %ld\n", what, i);
}

void fill(std::vector<Part> &P, std::vector<Cell> &C)
{
    for (size_t i = 0; i < P.size(); i++) {
        Part &p = P[i];
        Cell &c = C[i];
        p.Type = (short)(i % 7 == 3 ? 1 : 0);
        p.TimeBin = (short)(i % 40);
        p.wakeup = 0;
        for (int k = 0; k < 3; k++) {
            p.Pos[k] = rnd();
            c.VelPred[k] = rnd();
            c.HydroAccel[k] = rnd();
            c.Gradients.Density[k] = rnd();
            c.Gradients.Pressure[k] = rnd();
            c.NV_T_face_weights[k] = rnd();
            for (int d = 0; d < 3; d++) c.Gradients.Velocity[k][d] = rnd();
        }
        for (int k = 0; k < 6; k++) c.NV_T.data[k] = rnd();
        p.Mass = rnd();
        p.KernelRadius = (float)rnd();
        p.NumNgb = (float)rnd();
        p.DrkernNgbFactor = (float)rnd();
        p.Particle_DivVel = (float)rnd();
        c.Density = rnd();
        c.Pressure = rnd();
        c.InternalEnergyPred = rnd();
        c.DtInternalEnergy = rnd();
        c.ConditionNumber = rnd();
        c.FaceClosureError = rnd();
        c.MaxSignalVel = (float)rnd();
    }
}

bool same_bytes(const void *a, const void *b, size_t s) { return std::memcmp(a, b, s) == 0; }

}  // namespace

int main()
{
    const long n = 1000;
    std::vector<Part> P(n);
    std::vector<Cell> C(n);
    fill(P, C);
    gz_dropin_buffers b;
    gz_dropin_pack(b, P.data(), C.data(), n);
    const gz_fields &f = b.f;
    for (long i = 0; i < n; i++) {
        const Part &p = P[i];
        const Cell &c = C[i];
        for (int k = 0; k < 3; k++) {
            expect(f.pos[3 * i + k] == p.Pos[k], "pos", i);
            expect(f.vel[3 * i + k] == c.VelPred[k], "vel", i);
            expect(f.accel[3 * i + k] == c.HydroAccel[k], "accel", i);
            expect(f.grad_rho[3 * i + k] == c.Gradients.Density[k], "grad_rho", i);
            expect(f.grad_p[3 * i + k] == c.Gradients.Pressure[k], "grad_p", i);
            expect(f.nvt_fw[3 * i + k] == c.NV_T_face_weights[k], "nvt_fw", i);
            for (int d = 0; d < 3; d++) expect(f.grad_v[9 * i + 3 * k + d] == c.Gradients.Velocity[k][d], "grad_v", i);
        }
        for (int k = 0; k < 6; k++) expect(f.nvt[6 * i + k] == c.NV_T.data[k], "nvt", i);
        expect(f.mass[i] == p.Mass, "mass", i);
        expect(f.h[i] == (double)p.KernelRadius, "h", i);
        expect(f.numngb[i] == (double)p.NumNgb, "numngb", i);
        expect(f.drkern[i] == (double)p.DrkernNgbFactor, "drkern", i);
        expect(f.divvel[i] == (double)p.Particle_DivVel, "divvel", i);
        expect(f.u[i] == c.InternalEnergyPred, "u", i);
        expect(f.rho[i] == c.Density, "rho", i);
        expect(f.press[i] == c.Pressure, "press", i);
        expect(f.cn[i] == c.ConditionNumber, "cn", i);
        expect(f.fce[i] == c.FaceClosureError, "fce", i);
        expect(f.dtu[i] == c.DtInternalEnergy, "dtu", i);
        expect(f.maxsig[i] == (double)c.MaxSignalVel, "maxsig", i);
        expect(f.timebin[i] == p.TimeBin, "timebin", i);
        expect(f.wakeup[i] == p.wakeup, "wakeup", i);
    }
    expect(f.vel_kick == nullptr && f.u_kick == nullptr && f.ti_current == nullptr, "loop fields left NULL", -1);

    // new values in the flat arrays (as the library's stages would write them), then each stage
    // back on its own into a fresh copy of the particles
    std::vector<Part> P1 = P;
    std::vector<Cell> C1 = C;
    fill(P1, C1);
    gz_dropin_buffers nb;
    gz_dropin_pack(nb, P1.data(), C1.data(), n);
    for (long i = 0; i < n; i++) nb.f.wakeup[i] = (i == 617) ? 3 : 0;
    const int stages[3] = {GZ_DROPIN_DENSITY, GZ_DROPIN_GRADIENTS, GZ_DROPIN_FORCE};
    for (int s : stages) {
        std::vector<Part> Q = P;
        std::vector<Cell> D = C;
        const int woken = gz_dropin_unpack(nb, Q.data(), D.data(), n, s);
        expect(woken == (s == GZ_DROPIN_FORCE ? 1 : 0), "woken flag", s);
        for (long i = 0; i < n; i++) {
            const Part &src = (s == GZ_DROPIN_DENSITY || s == GZ_DROPIN_FORCE) ? P1[i] : P[i];
            const Cell &q = C1[i];
            const Cell &o = C[i];
            const bool dens = s == GZ_DROPIN_DENSITY, grad = s == GZ_DROPIN_GRADIENTS, force = s == GZ_DROPIN_FORCE;
            // members of the stage carry the new values, the others keep the old ones
            expect(Q[i].KernelRadius == (dens ? src.KernelRadius : P[i].KernelRadius), "KernelRadius", i);
            expect(Q[i].NumNgb == (dens ? P1[i].NumNgb : P[i].NumNgb), "NumNgb", i);
            expect(Q[i].DrkernNgbFactor == (dens ? P1[i].DrkernNgbFactor : P[i].DrkernNgbFactor), "DrkernNgbFactor", i);
            expect(Q[i].Particle_DivVel == (dens ? P1[i].Particle_DivVel : P[i].Particle_DivVel), "Particle_DivVel", i);
            expect(D[i].Density == (dens ? q.Density : o.Density), "Density", i);
            expect(D[i].Pressure == (dens ? q.Pressure : o.Pressure), "Pressure", i);
            expect(D[i].ConditionNumber == (dens ? q.ConditionNumber : o.ConditionNumber), "ConditionNumber", i);
            expect(D[i].FaceClosureError == (dens ? q.FaceClosureError : o.FaceClosureError), "FaceClosureError", i);
            expect(same_bytes(&D[i].NV_T, dens ? &q.NV_T : &o.NV_T, sizeof(Sym)), "NV_T", i);
            expect(same_bytes(&D[i].NV_T_face_weights, dens ? &q.NV_T_face_weights : &o.NV_T_face_weights, sizeof(V3<double>)),
                   "NV_T_face_weights", i);
            expect(same_bytes(&D[i].Gradients, grad ? &q.Gradients : &o.Gradients, sizeof(o.Gradients)), "Gradients", i);
            expect(same_bytes(&D[i].HydroAccel, force ? &q.HydroAccel : &o.HydroAccel, sizeof(V3<double>)), "HydroAccel", i);
            expect(D[i].DtInternalEnergy == (force ? q.DtInternalEnergy : o.DtInternalEnergy), "DtInternalEnergy", i);
            expect(D[i].MaxSignalVel == (force ? q.MaxSignalVel : o.MaxSignalVel), "MaxSignalVel", i);
            expect(Q[i].wakeup == (force ? (i == 617 ? 3 : 0) : P[i].wakeup), "wakeup", i);
            // never written by any stage
            expect(same_bytes(&Q[i].Pos, &P[i].Pos, sizeof(V3<double>)) && Q[i].Mass == P[i].Mass && Q[i].Type == P[i].Type &&
                       Q[i].TimeBin == P[i].TimeBin,
                   "P members outside the stages", i);
            expect(same_bytes(&D[i].VelPred, &o.VelPred, sizeof(V3<double>)) && D[i].InternalEnergyPred == o.InternalEnergyPred,
                   "CellP members outside the stages", i);
        }
    }

    // active flags: gas particles of the list only (Type 1 every 7th particle from 3 on), indices < n
    std::vector<int> list = {0, 3, 5, 10, 17, 999, 1000, 2000};
    const unsigned char *act = gz_dropin_active(b, P.data(), n, list);
    long on = 0;
    for (long i = 0; i < n; i++) {
        const bool want = (i == 0 || i == 5 || i == 999);  // 3, 10 and 17 are Type 1
        expect((act[i] != 0) == want, "active flag", i);
        on += act[i] != 0;
    }
    expect(on == 3, "active count", on);
    if (fails) {
        std::printf("dropin: %d mismatches\n", fails);
        return 1;
    }
    std::printf("dropin ok: %ld particles\n", n);
    return 0;
}
