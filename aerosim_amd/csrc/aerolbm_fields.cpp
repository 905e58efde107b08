// C ABI of the derived-field readouts (include/aerolbm.h): slice colour map, Q-criterion / vorticity and
// point probes, computed on the GPU from the context's resident macro fields (or, for aero_qcriterion, from
// fields the caller supplies).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerolbm.h"
#include "aero_ctx.h"
#include "fields.h"
#include "geometry.h"

using namespace aero;

namespace {

FieldView view_of(const aero_ctx* c) {
  const size_t N = c->N;
  return FieldView{c->mac, c->mac + N, c->mac + 2 * N, c->mac + 3 * N, c->flags, &c->st->corr, c->nx, c->ny, c->nz};
}

// max(0, min(n - 1, floor(fraction * n))) in double, as lbmCpuReference.ts:524-526
uint32_t slice_index(double fraction, uint32_t n) {
  double v = std::floor(fraction * n);
  if (v > n - 1.0) v = n - 1.0;
  if (v < 0) v = 0;
  return (uint32_t)v;
}

}  // namespace

extern "C" {

int aero_read_slice_rgba(aero_ctx* c, int32_t axis, double fraction, int32_t field, int32_t rules, double u_ref,
                         uint8_t* rgba, int32_t* wh) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, axis >= 0 && axis <= 2 && field >= 0 && field <= 2 && (rules == 0 || rules == 1) &&
                      std::isfinite(fraction) && !std::isnan(u_ref),
               AERO_E_ARG, "aero_read_slice_rgba: axis, field, rules or fraction out of range");
  const uint32_t w = axis == AERO_SLICE_X ? c->nz : c->nx;
  const uint32_t h = axis == AERO_SLICE_Y ? c->nz : c->ny;
  if (wh) {
    wh[0] = (int32_t)w;
    wh[1] = (int32_t)h;
  }
  if (!rgba) return AERO_OK;
  AERO_REQUIRE(c, c->readout_valid, AERO_E_STATE, "aero_read_slice_rgba: last aero_step wrote no macro fields");
  AERO_HIP(c, hipSetDevice(c->dev));
  if (!c->slice_buf) {
    const size_t m = (size_t)std::max(c->nx, c->nz) * std::max(c->ny, c->nz);
    AERO_HIP(c, hipMalloc(&c->slice_buf, m * 4));
  }
  const uint32_t n_axis = axis == AERO_SLICE_Z ? c->nz : axis == AERO_SLICE_Y ? c->ny : c->nx;
  const double ur = u_ref < 0 ? c->p.u_inlet : u_ref;
  AERO_HIP(c, launch_slice_rgba(view_of(c), axis, slice_index(fraction, n_axis), field, rules, ur, c->slice_buf,
                                c->stream));
  AERO_HIP(c, hipMemcpyAsync(rgba, c->slice_buf, (size_t)w * h * 4, hipMemcpyDeviceToHost, c->stream));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  return AERO_OK;
}

int aero_read_qcriterion(aero_ctx* c, double dx, float* q, float* omega_mag, float* omega_vec, double* maxima) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, std::isfinite(dx) && dx != 0, AERO_E_ARG, "aero_read_qcriterion: dx must be finite and non-zero");
  AERO_REQUIRE(c, c->readout_valid, AERO_E_STATE, "aero_read_qcriterion: last aero_step wrote no macro fields");
  AERO_HIP(c, hipSetDevice(c->dev));
  const size_t N = c->N;
  const FieldView v = view_of(c);
  if (!c->qbuf) AERO_HIP(c, hipMalloc(&c->qbuf, 5 * N * 4));
  if (!c->qpart) AERO_HIP(c, hipMalloc(&c->qpart, (2 * (size_t)qcrit_blocks(v) + 2) * 8));
  double* out = c->qpart + 2 * (size_t)qcrit_blocks(v);
  AERO_HIP(c, launch_qcrit(v, 1.0 / (1 * dx), 1.0 / (2.0 * dx), c->qbuf, c->qbuf + N, c->qbuf + 2 * N, c->qpart,
                           out, c->stream));
  if (q) AERO_HIP(c, hipMemcpyAsync(q, c->qbuf, N * 4, hipMemcpyDeviceToHost, c->stream));
  if (omega_mag) AERO_HIP(c, hipMemcpyAsync(omega_mag, c->qbuf + N, N * 4, hipMemcpyDeviceToHost, c->stream));
  if (omega_vec) AERO_HIP(c, hipMemcpyAsync(omega_vec, c->qbuf + 2 * N, 3 * N * 4, hipMemcpyDeviceToHost, c->stream));
  if (maxima) AERO_HIP(c, hipMemcpyAsync(maxima, out, 16, hipMemcpyDeviceToHost, c->stream));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  return AERO_OK;
}

int aero_qcriterion(int32_t device, int32_t nx, int32_t ny, int32_t nz, double dx, const float* ux, const float* uy,
                    const float* uz, const uint8_t* cell_flags, float* q, float* omega_mag, float* omega_vec,
                    double* maxima) {
  if (nx < 1 || ny < 1 || nz < 1 || !ux || !uy || !uz || !std::isfinite(dx) || dx == 0 ||
      (double)nx * ny * nz >= 2147483648.0) {
    set_host_error("aero_qcriterion: bad size, dx or null field");
    return AERO_E_ARG;
  }
  const size_t N = (size_t)nx * ny * nz;
  std::vector<void*> held;
  auto fail = [&](const char* what, hipError_t e) {
    for (void* p : held) hipFree(p);
    set_host_error(std::string("aero_qcriterion: ") + what + ": " + hipGetErrorString(e));
    return AERO_E_HIP;
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return fail("hipSetDevice", e);
  float* d = nullptr;
  if ((e = hipMalloc(&d, 8 * N * 4)) != hipSuccess) return fail("fields", e);
  held.push_back(d);
  uint8_t* fl = nullptr;
  if (cell_flags) {
    if ((e = hipMalloc(&fl, N)) != hipSuccess) return fail("flags", e);
    held.push_back(fl);
  }
  FieldView v{d, d + N, d + 2 * N, nullptr, fl, nullptr, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz};
  double* part = nullptr;
  if ((e = hipMalloc(&part, (2 * (size_t)qcrit_blocks(v) + 2) * 8)) != hipSuccess) return fail("partials", e);
  held.push_back(part);
  double* out = part + 2 * (size_t)qcrit_blocks(v);
  const float* src[3] = {ux, uy, uz};
  for (int k = 0; k < 3; ++k)
    if ((e = hipMemcpy(d + k * N, src[k], N * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail("upload", e);
  if (fl && (e = hipMemcpy(fl, cell_flags, N, hipMemcpyHostToDevice)) != hipSuccess) return fail("upload", e);
  float* o = d + 3 * N;
  if ((e = launch_qcrit(v, 1.0 / (1 * dx), 1.0 / (2.0 * dx), o, o + N, o + 2 * N, part, out, nullptr)) != hipSuccess)
    return fail("launch", e);
  if (q && (e = hipMemcpy(q, o, N * 4, hipMemcpyDeviceToHost)) != hipSuccess) return fail("download", e);
  if (omega_mag && (e = hipMemcpy(omega_mag, o + N, N * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return fail("download", e);
  if (omega_vec && (e = hipMemcpy(omega_vec, o + 2 * N, 3 * N * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return fail("download", e);
  if (maxima && (e = hipMemcpy(maxima, out, 16, hipMemcpyDeviceToHost)) != hipSuccess) return fail("download", e);
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail("synchronize", e);
  for (void* p : held) hipFree(p);
  return AERO_OK;
}

int aero_probe(aero_ctx* c, const aero_probe_params* p, const double* xyz, const uint8_t* on_surface, int32_t n,
               double* out) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, p && n >= 0 && (n == 0 || (xyz && out)), AERO_E_ARG, "aero_probe: null argument");
  AERO_REQUIRE(c, std::isfinite(p->dx) && p->dx != 0 && std::isfinite(p->origin[0]) && std::isfinite(p->origin[1]) &&
                      std::isfinite(p->origin[2]),
               AERO_E_ARG, "aero_probe: origin and dx must be finite, dx non-zero");
  for (int32_t i = 0; i < 3 * n; ++i)
    AERO_REQUIRE(c, std::isfinite(xyz[i]), AERO_E_ARG, "aero_probe: probe points must be finite");
  AERO_REQUIRE(c, c->readout_valid, AERO_E_STATE, "aero_probe: last aero_step wrote no macro fields");
  if (n == 0) return AERO_OK;
  AERO_HIP(c, hipSetDevice(c->dev));
  // per point: 3 doubles in, 13 out, 1 surface byte (rounded up to whole doubles)
  const size_t need = (size_t)n * 16 + ((size_t)n + 7) / 8;
  if (c->probe_cap < need) {
    if (c->probe_buf) AERO_HIP(c, hipFree(c->probe_buf));
    c->probe_buf = nullptr;
    c->probe_cap = 0;
    AERO_HIP(c, hipMalloc(&c->probe_buf, need * 8));
    c->probe_cap = need;
  }
  double* din = c->probe_buf;
  double* dout = din + 3 * (size_t)n;
  uint8_t* dsurf = on_surface ? (uint8_t*)(dout + 13 * (size_t)n) : nullptr;
  AERO_HIP(c, hipMemcpyAsync(din, xyz, (size_t)n * 24, hipMemcpyHostToDevice, c->stream));
  if (dsurf) AERO_HIP(c, hipMemcpyAsync(dsurf, on_surface, (size_t)n, hipMemcpyHostToDevice, c->stream));
  ProbeParams pp;
  for (int k = 0; k < 3; ++k) pp.origin[k] = p->origin[k];
  pp.dx = p->dx;
  pp.wind_speed_mps = p->wind_speed_mps;
  pp.u_inlet_lattice = p->u_inlet_lattice;
  pp.fluid_density = p->fluid_density;
  AERO_HIP(c, launch_probe(view_of(c), pp, din, dsurf, (uint32_t)n, dout, c->stream));
  AERO_HIP(c, hipMemcpyAsync(out, dout, (size_t)n * 13 * 8, hipMemcpyDeviceToHost, c->stream));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  return AERO_OK;
}

int aero_mass_stats(aero_ctx* c, double* out) {
  if (!c || !out) return AERO_E_ARG;
  AERO_REQUIRE(c, c->readout_valid, AERO_E_STATE, "aero_mass_stats: last aero_step wrote no macro fields");
  AERO_HIP(c, hipSetDevice(c->dev));
  const uint32_t nb = mass_stats_blocks(c->N);
  if (!c->mass_buf) AERO_HIP(c, hipMalloc(&c->mass_buf, ((size_t)4 * nb + 8) * 8));
  double* res = c->mass_buf + (size_t)4 * nb;
  AERO_HIP(c, launch_mass_stats(view_of(c), c->mass_buf, res, c->stream));
  AERO_HIP(c, hipMemcpyAsync(out, res, 5 * 8, hipMemcpyDeviceToHost, c->stream));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  return AERO_OK;
}

}  // extern "C"
