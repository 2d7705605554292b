/*
 * gizmo_dropin.h -- host-side helpers of the drop-in shim (C++; header only, no device code).
 *
 * The shim that replaces GIZMO's density(), hydro_gradient_calc() and hydro_force()
 * (integration/gz_dropin.cc) moves the hot-path fields of the gas particles between the
 * reference's particle arrays P[] (struct particle_data, src/declarations/particle_data.h) and
 * CellP[] (struct cell_data, src/declarations/cell_data.h) and the flat arrays of gz_fields
 * (include/gizmo_hydro.h).  The copies are templates over the two element types, so they
 * compile against any type with the members named below (Vec3-like members indexed [k],
 * Gradients.Velocity indexed [k][d], NV_T.data[6]): GIZMO's own structs, or the test's.
 *
 *   gz_dropin_pack      P[]/CellP[] -> flat arrays; fills a gz_fields with their pointers
 *   gz_dropin_unpack    the fields a stage writes, flat arrays -> P[]/CellP[]
 *   gz_dropin_active    the active flags of the gas particles of an active-particle list
 *
 * Fields per stage (what unpack writes): density() writes KernelRadius, NumNgb,
 * DrkernNgbFactor, Particle_DivVel, Density, Pressure, NV_T, NV_T_face_weights,
 * ConditionNumber, FaceClosureError; hydro_gradient_calc() writes Gradients.{Density,
 * Pressure, Velocity}; hydro_force() writes HydroAccel, DtInternalEnergy, MaxSignalVel and
 * the wakeup flags of inactive neighbours.
 */
#ifndef GIZMO_DROPIN_H
#define GIZMO_DROPIN_H

#include <vector>

#include "gizmo_hydro.h"

/* which P[]/CellP[] fields gz_dropin_unpack writes back */
enum {
    GZ_DROPIN_DENSITY = 1,
    GZ_DROPIN_GRADIENTS = 2,
    GZ_DROPIN_FORCE = 4,
    GZ_DROPIN_ALL = 7,
};

/* host copies of the gas particles' hot-path fields, kept between calls (no reallocation once
   they hold the largest N seen) */
struct gz_dropin_buffers {
    std::vector<double> pos, vel, mass, h, u, numngb, drkern, divvel, rho, press, nvt, nvt_fw, cn, fce, grad_rho, grad_p,
        grad_v, accel, dtu, maxsig;
    std::vector<int> timebin, wakeup;
    std::vector<unsigned char> active;
    gz_fields f{};
    long n = 0;

    void resize(long n_)
    {
        n = n_;
        const size_t s = (size_t)n_;
        auto grow = [s](std::vector<double> &v, int w) {
            v.resize(s * w);
            return v.data();
        };
        f = gz_fields{};
        f.pos = grow(pos, 3); f.vel = grow(vel, 3); f.mass = grow(mass, 1); f.h = grow(h, 1); f.u = grow(u, 1);
        f.numngb = grow(numngb, 1); f.drkern = grow(drkern, 1); f.divvel = grow(divvel, 1); f.rho = grow(rho, 1);
        f.press = grow(press, 1); f.nvt = grow(nvt, 6); f.nvt_fw = grow(nvt_fw, 3); f.cn = grow(cn, 1);
        f.fce = grow(fce, 1); f.grad_rho = grow(grad_rho, 3); f.grad_p = grow(grad_p, 3); f.grad_v = grow(grad_v, 9);
        f.accel = grow(accel, 3); f.dtu = grow(dtu, 1); f.maxsig = grow(maxsig, 1);
        timebin.resize(s); wakeup.resize(s);
        f.timebin = timebin.data(); f.wakeup = wakeup.data();
    }
};

/* P[0..n-1] / CellP[0..n-1] (the gas particles) into b, and b.f pointing at its arrays */
template <class PartT, class CellT>
void gz_dropin_pack(gz_dropin_buffers &b, const PartT *P, const CellT *CellP, long n)
{
    b.resize(n);
    gz_fields &f = b.f;
    for (long i = 0; i < n; i++) {
        const PartT &p = P[i];
        const CellT &c = CellP[i];
        for (int k = 0; k < 3; k++) {
            f.pos[3 * i + k] = p.Pos[k];
            f.vel[3 * i + k] = c.VelPred[k];
            f.nvt_fw[3 * i + k] = c.NV_T_face_weights[k];
            f.grad_rho[3 * i + k] = c.Gradients.Density[k];
            f.grad_p[3 * i + k] = c.Gradients.Pressure[k];
            f.accel[3 * i + k] = c.HydroAccel[k];
            for (int d = 0; d < 3; d++) f.grad_v[9 * i + 3 * k + d] = c.Gradients.Velocity[k][d];
        }
        for (int k = 0; k < 6; k++) f.nvt[6 * i + k] = c.NV_T.data[k];
        f.mass[i] = p.Mass;
        f.h[i] = p.KernelRadius;
        f.numngb[i] = p.NumNgb;
        f.drkern[i] = p.DrkernNgbFactor;
        f.divvel[i] = p.Particle_DivVel;
        f.u[i] = c.InternalEnergyPred;
        f.rho[i] = c.Density;
        f.press[i] = c.Pressure;
        f.cn[i] = c.ConditionNumber;
        f.fce[i] = c.FaceClosureError;
        f.dtu[i] = c.DtInternalEnergy;
        f.maxsig[i] = c.MaxSignalVel;
        f.timebin[i] = p.TimeBin;
        f.wakeup[i] = p.wakeup;
    }
}

/* the fields of the stages in `stages` (GZ_DROPIN_*) from b back into P[]/CellP[]; returns 1 when
   a particle carries a wakeup flag (the caller's NeedToWakeupParticles_local), else 0 */
template <class PartT, class CellT>
int gz_dropin_unpack(const gz_dropin_buffers &b, PartT *P, CellT *CellP, long n, int stages)
{
    const gz_fields &f = b.f;
    int woken = 0;
    for (long i = 0; i < n; i++) {
        PartT &p = P[i];
        CellT &c = CellP[i];
        if (stages & GZ_DROPIN_DENSITY) {
            p.KernelRadius = f.h[i];
            p.NumNgb = f.numngb[i];
            p.DrkernNgbFactor = f.drkern[i];
            p.Particle_DivVel = f.divvel[i];
            c.Density = f.rho[i];
            c.Pressure = f.press[i];
            c.ConditionNumber = f.cn[i];
            c.FaceClosureError = f.fce[i];
            for (int k = 0; k < 6; k++) c.NV_T.data[k] = f.nvt[6 * i + k];
            for (int k = 0; k < 3; k++) c.NV_T_face_weights[k] = f.nvt_fw[3 * i + k];
        }
        if (stages & GZ_DROPIN_GRADIENTS) {
            for (int k = 0; k < 3; k++) {
                c.Gradients.Density[k] = f.grad_rho[3 * i + k];
                c.Gradients.Pressure[k] = f.grad_p[3 * i + k];
                for (int d = 0; d < 3; d++) c.Gradients.Velocity[k][d] = f.grad_v[9 * i + 3 * k + d];
            }
        }
        if (stages & GZ_DROPIN_FORCE) {
            for (int k = 0; k < 3; k++) c.HydroAccel[k] = f.accel[3 * i + k];
            c.DtInternalEnergy = f.dtu[i];
            c.MaxSignalVel = f.maxsig[i];
            p.wakeup = f.wakeup[i];
            if (f.wakeup[i]) woken = 1;
        }
    }
    return woken;
}

/* the active flags gz_set_active takes: 1 for the gas particles (index < n, Type 0) of an
   active-particle list (the reference's ActiveParticleList, src/core/run.cc:997-1009) */
template <class PartT, class ListT>
const unsigned char *gz_dropin_active(gz_dropin_buffers &b, const PartT *P, long n, const ListT &active_list)
{
    b.active.assign((size_t)n, 0);
    for (auto i : active_list)
        if ((long)i < n && P[i].Type == 0) b.active[(size_t)i] = 1;
    return b.active.data();
}

#endif /* GIZMO_DROPIN_H */
