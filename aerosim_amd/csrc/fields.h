// Launchers of the derived-field readout kernels (defined in fields.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aero {

// Macro fields a readout reads: SoA float arrays of N cells, the cell flags, and the mass-anchor factor
// that the solver's density still owes (rho of a non-solid cell is (float)((double)rho * *corr); corr may
// be null for 1).
struct FieldView {
  const float *ux, *uy, *uz, *rho;
  const uint8_t* flags;  // may be null: no cell is solid
  const double* corr;
  uint32_t nx, ny, nz;
};

// Slice pixel rules: extractSliceRGBA of the CPU solver (FP64) or the WGSL post-process shader (FP32).
enum SliceRules { SLICE_RULES_CPU = 0, SLICE_RULES_WGSL = 1 };

// One RGBA8 pixel per cell of the plane axis = fixed (axis 0: z, 1: y, 2: x), written as width*height
// little-endian words (bytes R, G, B, A).
hipError_t launch_slice_rgba(const FieldView& v, int axis, uint32_t fixed, int field, int rules, double u_ref,
                             uint32_t* rgba, hipStream_t s);

// Q-criterion, vorticity magnitude and vorticity vector (3 floats per cell) of every cell; part receives
// 2 doubles (max Q, max |omega|, both at least 1e-6) per block of qcrit_blocks(); out receives the 2 maxima.
uint32_t qcrit_blocks(const FieldView& v);
hipError_t launch_qcrit(const FieldView& v, double inv_dx, double inv_2dx, float* q, float* om, float* ov,
                        double* part, double* out, hipStream_t s);

// computeFluidMassStats over the macro fields: part receives 4 doubles per block of mass_stats_blocks(N); out
// receives total mass, mean rho, largest speed, NaN flag and counted cells.
uint32_t mass_stats_blocks(uint32_t N);
hipError_t launch_mass_stats(const FieldView& v, double* part, double* out, hipStream_t s);

// Point probes: per point the world position (3 doubles) and whether the ray hit the mesh surface (u8);
// per point 13 doubles out (see aero_probe).
struct ProbeParams {
  double origin[3], dx;
  double wind_speed_mps, u_inlet_lattice, fluid_density;
};
hipError_t launch_probe(const FieldView& v, const ProbeParams& p, const double* xyz, const uint8_t* surface,
                        uint32_t n, double* out, hipStream_t s);

}  // namespace aero
