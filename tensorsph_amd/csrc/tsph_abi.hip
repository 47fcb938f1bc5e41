// Layout of the parameter structs of the C ABI, so a binding that restates
// them (ctypes, cgo, JNI) can check its copy against the compiled one.
#include <stddef.h>

#include "tsph_internal.hpp"

namespace {

struct FieldInfo {
  const char* st;
  const char* field;
  int64_t offset, size, struct_size;
};

#define TSPH_F(S, F) {#S, #F, (int64_t)offsetof(S, F), (int64_t)sizeof(((S*)0)->F), (int64_t)sizeof(S)}

const FieldInfo kFields[] = {
    TSPH_F(tsph_wcsph_params, dim), TSPH_F(tsph_wcsph_params, precision),
    TSPH_F(tsph_wcsph_params, n_fluid), TSPH_F(tsph_wcsph_params, n_wall), TSPH_F(tsph_wcsph_params, h),
    TSPH_F(tsph_wcsph_params, rho0), TSPH_F(tsph_wcsph_params, sound_speed), TSPH_F(tsph_wcsph_params, gamma),
    TSPH_F(tsph_wcsph_params, gravity), TSPH_F(tsph_wcsph_params, viscosity),
    TSPH_F(tsph_wcsph_params, wall_friction), TSPH_F(tsph_wcsph_params, dt),
    TSPH_F(tsph_wcsph_params, diffusion_scale), TSPH_F(tsph_wcsph_params, diffusion_form),
    TSPH_F(tsph_wcsph_params, eigenvalue_floor), TSPH_F(tsph_wcsph_params, reaction_style),
    TSPH_F(tsph_wcsph_params, has_wall_accel), TSPH_F(tsph_wcsph_params, wrap_width),
    TSPH_F(tsph_wcsph_params, density_renormalization), TSPH_F(tsph_wcsph_params, lattice_kernel_sum),
    TSPH_F(tsph_wcsph_params, boundary_consistent_density),
    TSPH_F(tsph_wcsph_params, surface_pressure_closure),
    TSPH_F(tsph_wcsph_params, surface_pressure_threshold),
    TSPH_F(tsph_wcsph_params, surface_pressure_donors_all),
    TSPH_F(tsph_wcsph_params, surface_pressure_layers),

    TSPH_F(tsph_threed_params, precision), TSPH_F(tsph_threed_params, n), TSPH_F(tsph_threed_params, n_fluid),
    TSPH_F(tsph_threed_params, h), TSPH_F(tsph_threed_params, rho0), TSPH_F(tsph_threed_params, sound_speed),
    TSPH_F(tsph_threed_params, gamma), TSPH_F(tsph_threed_params, viscosity),
    TSPH_F(tsph_threed_params, gravity), TSPH_F(tsph_threed_params, has_gravity),
    TSPH_F(tsph_threed_params, box), TSPH_F(tsph_threed_params, periodic),
    TSPH_F(tsph_threed_params, renormalise), TSPH_F(tsph_threed_params, diffusion_scale),
    TSPH_F(tsph_threed_params, diffusion_form), TSPH_F(tsph_threed_params, eigenvalue_floor),
    TSPH_F(tsph_threed_params, dt), TSPH_F(tsph_threed_params, wall_model),
    TSPH_F(tsph_threed_params, viscous_order),

    TSPH_F(tsph_dp_params, precision), TSPH_F(tsph_dp_params, n), TSPH_F(tsph_dp_params, n_wall),
    TSPH_F(tsph_dp_params, dx), TSPH_F(tsph_dp_params, sound_speed), TSPH_F(tsph_dp_params, reference_speed),
    TSPH_F(tsph_dp_params, dt), TSPH_F(tsph_dp_params, rho0), TSPH_F(tsph_dp_params, gravity),
    TSPH_F(tsph_dp_params, viscosity), TSPH_F(tsph_dp_params, diffusion), TSPH_F(tsph_dp_params, alpha),

    TSPH_F(tsph_batch_caps, fluid), TSPH_F(tsph_batch_caps, wall), TSPH_F(tsph_batch_caps, sims),
    TSPH_F(tsph_batch_caps, cells), TSPH_F(tsph_batch_caps, fluid_entries),
    TSPH_F(tsph_batch_caps, wall_entries), TSPH_F(tsph_batch_caps, sorted_lists),

    TSPH_F(tsph_dp_case, n), TSPH_F(tsph_dp_case, n_wall), TSPH_F(tsph_dp_case, steps),
    TSPH_F(tsph_dp_case, dx), TSPH_F(tsph_dp_case, sound_speed), TSPH_F(tsph_dp_case, reference_speed),
    TSPH_F(tsph_dp_case, rho0), TSPH_F(tsph_dp_case, gravity), TSPH_F(tsph_dp_case, viscosity),
    TSPH_F(tsph_dp_case, diffusion), TSPH_F(tsph_dp_case, alpha), TSPH_F(tsph_dp_case, courant),

    TSPH_F(tsph_sim_params, op), TSPH_F(tsph_sim_params, skin), TSPH_F(tsph_sim_params, lo),
    TSPH_F(tsph_sim_params, hi), TSPH_F(tsph_sim_params, tol), TSPH_F(tsph_sim_params, clamp),
    TSPH_F(tsph_sim_params, obstacle_begin), TSPH_F(tsph_sim_params, obstacle_end),
    TSPH_F(tsph_sim_params, obstacle_centre), TSPH_F(tsph_sim_params, obstacle_clamp),
    TSPH_F(tsph_sim_params, obstacle_lo), TSPH_F(tsph_sim_params, obstacle_hi),
    TSPH_F(tsph_sim_params, obstacle_eps), TSPH_F(tsph_sim_params, adaptive),
    TSPH_F(tsph_sim_params, target_time), TSPH_F(tsph_sim_params, courant_num),
    TSPH_F(tsph_sim_params, dt_limit), TSPH_F(tsph_sim_params, dt_safety), TSPH_F(tsph_sim_params, dt_growth),
    TSPH_F(tsph_sim_params, adaptive_free_last), TSPH_F(tsph_sim_params, adaptive_tol),
    TSPH_F(tsph_sim_params, initial_time), TSPH_F(tsph_sim_params, initial_dt),

    TSPH_F(tsph_shift_params, dt), TSPH_F(tsph_shift_params, v_ref), TSPH_F(tsph_shift_params, v_floor),
    TSPH_F(tsph_shift_params, strength), TSPH_F(tsph_shift_params, limit),
    TSPH_F(tsph_shift_params, free_surface), TSPH_F(tsph_shift_params, percentile),
    TSPH_F(tsph_shift_params, c_reference),

    TSPH_F(tsph_tvc_params, alpha), TSPH_F(tsph_tvc_params, limiter_slope), TSPH_F(tsph_tvc_params, max_shift),
    TSPH_F(tsph_tvc_params, wall_volume), TSPH_F(tsph_tvc_params, scope), TSPH_F(tsph_tvc_params, bulk_c),
    TSPH_F(tsph_tvc_params, bulk_g), TSPH_F(tsph_tvc_params, percentile),

    TSPH_F(tsph_gauge_spec, station_x), TSPH_F(tsph_gauge_spec, aperture), TSPH_F(tsph_gauge_spec, tank_depth),
    TSPH_F(tsph_gauge_spec, dx), TSPH_F(tsph_gauge_spec, h), TSPH_F(tsph_gauge_spec, gravity),
    TSPH_F(tsph_gauge_spec, obstacle_lo), TSPH_F(tsph_gauge_spec, obstacle_hi),
    TSPH_F(tsph_gauge_spec, per_tap), TSPH_F(tsph_gauge_spec, nodes),
};

#undef TSPH_F

}  // namespace

extern "C" int tsph_abi_field(int index, const char** struct_name, const char** field_name, int64_t* offset,
                              int64_t* size, int64_t* struct_size) {
  const int count = (int)(sizeof(kFields) / sizeof(kFields[0]));
  if (index < 0 || index >= count) {
    tsph::set_error("field index %d outside [0, %d)", index, count);
    return TSPH_ERR_ARG;
  }
  const FieldInfo& f = kFields[index];
  if (struct_name) *struct_name = f.st;
  if (field_name) *field_name = f.field;
  if (offset) *offset = f.offset;
  if (size) *size = f.size;
  if (struct_size) *struct_size = f.struct_size;
  return TSPH_OK;
}
