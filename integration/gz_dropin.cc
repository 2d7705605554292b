// gz_dropin.cc -- the file a GIZMO maintainer adds to GIZMO's tree (as src/hydro/gz_dropin.cc)
// to run the three hot-path entry points on the MI355X through the C ABI of include/gizmo_hydro.h.
//
// It replaces
//   void density(void);              src/hydro/density_loop.cc:764
//   void hydro_gradient_calc(void);  src/hydro/gradients_loop.cc:458
//   void hydro_force(void);          src/hydro/hydro_force_loop.cc:423
// which compute_hydro_densities_and_forces() calls (src/core/accel.cc:80,143,169); the objects
// density_loop.o, gradients_loop.o and hydro_force_loop.o leave GIZMO's link line.  It needs
// GIZMO's headers (allvars.h, proto.h) and so compiles only inside GIZMO's tree; the copies
// between P[]/CellP[] and the library's flat arrays are the templates of include/gizmo_dropin.h,
// which tests/test_dropin.py compiles and checks on the CPU.
//
// Build (GIZMO side): -I<this repository>/include, and link
//   -L<this repository>/gizmo_amd -lgizmo_hip -Wl,-rpath,<this repository>/gizmo_amd
#include "../declarations/allvars.h"
#include "../core/proto.h"
#include "gizmo_dropin.h"

static gz_context *gz = nullptr;
static gz_dropin_buffers buf;

// the All parameters the three stages read (src/declarations/global_data_all_struct.h)
static gz_params params_from_All(void)
{
    gz_params p{};
    p.box_size = All.BoxSize;
    p.kernel = KERNEL_FUNCTION;
    p.gamma = GAMMA_DEFAULT;
    p.des_num_ngb = All.DesNumNgb;
    p.max_num_ngb_dev = All.MaxNumNgbDeviation;
    p.min_kernel_radius = All.MinKernelRadius;
    p.max_kernel_radius = All.MaxKernelRadius;
    p.time = All.Time;
    p.time_begin = All.TimeBegin;
    p.precision = GZ_FP64;
    return p;
}

// a nonzero code ends the run as the replaced stage would (endrun with the reference's code)
static void check(int rc)
{
    if (rc) {
        PRINT_WARNING("gizmo_hydro: %s", gz_last_error(gz));
        endrun(rc);
    }
}

static void store_back(int stages)
{
    check(gz_store(gz, &buf.f));
    if (gz_dropin_unpack(buf, P, CellP, N_gas, stages)) NeedToWakeupParticles_local = 1;
}

void density(void)
{
    gz_params p = params_from_All();
    if (!gz) {
        gz = gz_create(&p, All.MaxPartGas, 0);
        if (!gz) endrun(1);
    } else {
        check(gz_set_params(gz, &p));
    }
    // the particles moved since the last call: the load orders them again on the device
    gz_dropin_pack(buf, P, CellP, N_gas);
    check(gz_load(gz, N_gas, &buf.f));
    check(gz_set_active(gz, gz_dropin_active(buf, P, N_gas, ActiveParticleList)));
    check(gz_density(gz));
    store_back(GZ_DROPIN_DENSITY);
}

void hydro_gradient_calc(void)
{
    check(gz_hydro_gradient_calc(gz));
    store_back(GZ_DROPIN_GRADIENTS);
}

void hydro_force(void)
{
    check(gz_hydro_force(gz));
    store_back(GZ_DROPIN_FORCE);
}
