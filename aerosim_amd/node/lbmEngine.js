'use strict';
/**
 * JavaScript host of the MI355X D3Q19 wind-tunnel solver. Keeps the interface of the reference's
 * LBMEngine (src/solver/lbmEngine.ts:39-710) and voxelizeMeshToLattice (src/geometry/voxelizer.ts:86) and
 * calls the HIP library through the N-API addon (aerolbm.node). There is no CPU fallback: constructing an
 * engine without a usable GPU throws.
 */
const path = require('path');
const addon = require(path.join(__dirname, 'aerolbm.node'));

const DEFAULT_CONFIG = {
  uInletLattice: 0.085,
  nuLattice: 0.0005,
  smagorinskyCs: 0.12,
  groundNoSlip: false,
  windSpeedMps: 25.0,
  fluid: { density: 1.204 },
  frontalAreaM2: Math.PI * 0.25,
  characteristicLengthM: 1.0,
  tunnel: { length: 6.0, width: 2.8, height: 2.4 },
  sliceAxis: 'z',
  sliceFraction: 0.5,
  fieldMode: 'vorticity',
  sliceRules: 'wgsl', // the reference GPU engine's post-process shader; 'cpu' for its CPU solver's rules
};

// AERO_STEP_MACROS of include/aerolbm.h: the last step writes the macro fields only
const STEP_MACROS = 2;
const AXIS_CODE = { z: 0, y: 1, x: 2 };
const FIELD_CODE = { velocity: 0, vorticity: 1, cp: 2 };

const NEWTONS_TO_LBF = 0.224808943; // lbmConstants.ts:101-102
const NEWTONS_TO_KGF = 0.101971621;

/** Rolling force statistics (forceIntegrator.ts:321-424). */
class ForceConvergenceTracker {
  constructor() { this.reset(60); }
  reset(windowSize = 60) { this.windowSize = windowSize; this.samples = []; }
  pushSample(s) {
    this.samples.push({ step: s.step, cd: s.cd, cl: s.cl, cm: s.cm, xcp: s.xcp });
    if (this.samples.length > 180) this.samples.shift();
    const recent = this.samples.slice(-this.windowSize);
    const n = recent.length;
    const sum = { cd: 0, cl: 0, cm: 0, xcp: 0 };
    for (const r of recent) { sum.cd += r.cd; sum.cl += r.cl; sum.cm += r.cm; sum.xcp += r.xcp; }
    const meanCd = sum.cd / n;
    const meanCl = sum.cl / n;
    let varCd = 0;
    for (const r of recent) varCd += (r.cd - meanCd) * (r.cd - meanCd);
    const stdCd = Math.sqrt(varCd / Math.max(1, n));
    const relStd = stdCd / Math.max(0.02, Math.abs(meanCd));
    const fill = Math.min(1.0, n / 25);
    const stabilityPercent = Number(Math.max(0, Math.min(99.9, fill * (1.0 - Math.min(1.0, relStd * 4.0)) * 100)).toFixed(1));
    const qArea = 0.5 * s.fluid.density * s.windSpeed * s.windSpeed * s.frontalAreaM2;
    const drag = qArea * meanCd;
    const lift = qArea * meanCl;
    return {
      instantCd: s.cd, instantCl: s.cl, meanCd, meanCl, stdCd,
      dragForceNewtons: drag, dragForceLbs: drag * NEWTONS_TO_LBF, dragForceKgf: drag * NEWTONS_TO_KGF,
      liftForceNewtons: lift, liftForceLbs: lift * NEWTONS_TO_LBF,
      pitchingMomentCm: sum.cm / n, centerOfPressureX: sum.xcp / n,
      stabilityPercent, isConverged: n >= 20 && stabilityPercent >= 96.0,
      history: this.samples.map((r) => ({ step: r.step, cd: r.cd, cl: r.cl, cm: r.cm, xcp: r.xcp,
        dragForceN: r.cd * qArea, liftForceN: r.cl * qArea })),
    };
  }
}

const pick = (...v) => { for (const x of v) if (x !== undefined && x !== null) return x; return undefined; };

/** exportTelemetryToCSV (domainWorkflows.ts:323-371): RFC 4180 CSV with a comment header, CRLF line ends. */
function exportTelemetryToCSV(history, metadata) {
  const m = metadata || {};
  const fx = (v, d, dflt) => (v !== undefined && v !== null ? v.toFixed(d) : dflt);
  const lines = [
    '# AeroSim Aerodynamic Wind Tunnel Telemetry Export',
    `# Generated: ${pick(m.timestamp, new Date().toISOString())}`,
    `# Specimen: ${pick(m.specimenName, 'Specimen')}`,
    `# Wind Speed: ${fx(m.windSpeedMps, 1, '0.0')} m/s`,
    `# Reynolds Number: ${fx(m.reynolds, 0, '0')}`,
    `# Frontal Area: ${fx(m.meshStats ? m.meshStats.frontalArea : undefined, 4, '0')} m2`,
    `# Mean Cd: ${fx(m.meanCd, 4, '0')}`,
    `# Mean Cl: ${fx(m.meanCl, 4, '0')}`,
    `# Pitching Moment Cm: ${fx(m.meanCm, 4, '0')}`,
    `# Center of Pressure X: ${fx(m.centerOfPressureX, 3, '0')} m`,
    '',
    'step,time_sec,cd,cl,cm,drag_force_n,drag_force_lbf,drag_force_kgf,lift_force_n,lift_force_lbf,xcp_meters',
  ];
  for (const s of history) {
    const drag = pick(s.dragForceN, m.dragForceN, 0.0);
    const lift = pick(s.liftForceN, m.liftForceN, 0.0);
    lines.push([s.step, s.timeSec.toFixed(4), s.cd.toFixed(5), s.cl.toFixed(5), pick(s.cm, m.meanCm, 0.0).toFixed(5),
      drag.toFixed(2), (drag * NEWTONS_TO_LBF).toFixed(2), (drag * NEWTONS_TO_KGF).toFixed(2), lift.toFixed(2),
      (lift * NEWTONS_TO_LBF).toFixed(2), pick(s.xcpMeters, m.centerOfPressureX, 0.0).toFixed(3)].join(','));
  }
  return lines.join('\r\n');
}

/** exportRunReportToJSON (domainWorkflows.ts:376-402), with the reference report's producer name and version. */
function exportRunReportToJSON(metadata, history, exportTimestamp,
  application = 'AeroSim \u2014 3D WebGPU Wind Tunnel Platform', version = '0.1.0') {
  const r = (v, d) => (v !== undefined && v !== null ? Number(v.toFixed(d)) : undefined);
  const run = Object.assign({}, metadata, {
    units: { windSpeed: 'm/s', forces: 'Newtons', pressure: 'Pascals', lengths: 'meters', areas: 'square_meters' },
    telemetryHistory: history.map((s) => ({
      step: s.step, timeSec: Number(s.timeSec.toFixed(4)), cd: Number(s.cd.toFixed(5)), cl: Number(s.cl.toFixed(5)),
      cm: r(s.cm, 5), dragForceN: r(s.dragForceN, 2), liftForceN: r(s.liftForceN, 2), xcpMeters: r(s.xcpMeters, 3),
    })),
  });
  return JSON.stringify({ application, version, exportTimestamp: exportTimestamp || new Date().toISOString(),
    simulationRun: run }, null, 2);
}

/** computeAerodynamicTelemetry (fluidProperties.ts:7-52). */
function computeAerodynamicTelemetry({ windSpeed, characteristicLength, frontalArea, fluid, tunnel }) {
  const speed = Math.max(0, windSpeed);
  const re = (speed * Math.max(1e-6, characteristicLength)) / Math.max(1e-12, fluid.kinematicViscosity);
  let flowRegime = 'Laminar';
  if (re >= 3.5e5) flowRegime = 'Turbulent (Supercritical)';
  else if (re >= 2000) flowRegime = 'Turbulent (Subcritical)';
  else if (re >= 500) flowRegime = 'Transitional';
  return {
    reynoldsNumber: re, machNumber: speed / Math.max(1, fluid.speedOfSound),
    dynamicPressurePa: 0.5 * fluid.density * speed * speed, frontalAreaM2: frontalArea,
    characteristicLengthM: characteristicLength,
    blockageRatioPercent: (frontalArea / Math.max(1e-4, tunnel.width * tunnel.height)) * 100, flowRegime,
  };
}

/** voxelizeMeshToLattice on the GPU. geometry: {position: Float32Array} (non-indexed triangle soup). */
function voxelizeMeshToLattice(geometry, worldMatrix, tunnel, grid, device = 0) {
  const ctx = addon.createContext(grid.nx, grid.ny, grid.nz, 0, device);
  try {
    const world = worldMatrix ? Float64Array.from(worldMatrix.elements || worldMatrix) : null;
    addon.setMesh(ctx, geometry.position, world, tunnel);
    const n = grid.nx * grid.ny * grid.nz;
    const counts = addon.voxelCounts(ctx);
    const arrays = addon.readVoxelArrays(ctx, n);
    const h = tunnel.height / grid.ny;
    return Object.assign({
      nx: grid.nx, ny: grid.ny, nz: grid.nz, dx: h, dy: h, dz: h,
      originX: -grid.nx * h * 0.5, originY: -grid.ny * h * 0.5, originZ: -grid.nz * h * 0.5,
    }, arrays, counts);
  } finally {
    addon.destroy(ctx);
  }
}

class LBMEngine {
  /** device: HIP ordinal (the reference passes a GPUDevice; null there selects its CPU solver, which this
   * build does not have). */
  constructor(device = 0, storage = 0) {
    if (device === null || device === undefined) throw new Error('LBMEngine: a HIP device ordinal is required');
    this.device = device;
    this.storage = storage;
    this.ctx = null;
    this.voxel = null;
    this.dims = null;
    this.config = Object.assign({}, DEFAULT_CONFIG);
    this.stepCount = 0;
    this.isBusy = false;
    this.forceTracker = new ForceConvergenceTracker();
  }

  setVoxelGrid(voxel, config = {}, resetFlow = true) {
    const { nx, ny, nz } = voxel;
    if (!this.ctx || this.dims.nx !== nx || this.dims.ny !== ny || this.dims.nz !== nz) {
      if (this.ctx) addon.destroy(this.ctx);
      this.ctx = addon.createContext(nx, ny, nz, this.storage, this.device);
      this.dims = { nx, ny, nz };
    }
    this.voxel = voxel;
    addon.setVoxels(this.ctx, voxel.cellFlags, voxel.boundaryIndexMap, voxel.boundaryCellIndices, voxel.boundaryQValues);
    this.updateConfig(config);
    if (resetFlow) this.resetFlow();
  }

  updateConfig(partial) {
    Object.assign(this.config, partial);
    if (this.ctx) {
      const c = this.config;
      addon.setParams(this.ctx, { uInlet: c.uInletLattice, nu: c.nuLattice, smagorinskyCs: c.smagorinskyCs,
        groundNoSlip: !!c.groundNoSlip });
    }
  }

  resetFlow() {
    if (!this.ctx) return;
    addon.reset(this.ctx, 0, 0.0);
    this.stepCount = 0;
    this.forceTracker.reset(60);
  }

  getStepCount() { return this.stepCount; }

  /** stepAndExtractSlice (lbmEngine.ts:404): substeps on the GPU, then the force readout. Resolves null when
   * busy or without geometry, like the reference. */
  async stepAndExtractSlice(substeps = 4, readMacroField = false) {
    if (!this.ctx || !this.voxel || this.isBusy) return null;
    this.isBusy = true;
    try {
      const c = this.config;
      // the last substep writes the macro fields the slice reads (not the population snapshot)
      this.stepCount = Number(addon.step(this.ctx, substeps, STEP_MACROS));
      const f = addon.readForces(this.ctx, {
        frontalAreaM2: c.frontalAreaM2, characteristicLengthM: c.characteristicLengthM,
        tunnelWidth: c.tunnel.width, tunnelHeight: c.tunnel.height, dx: this.voxel.dx || (c.tunnel.height / this.voxel.ny),
        uInletLattice: c.uInletLattice,
      });
      const forceReadout = this.forceTracker.pushSample({
        step: this.stepCount, cd: f.cd, cl: f.cl, cm: f.cm, xcp: f.xcp, windSpeed: c.windSpeedMps,
        fluid: c.fluid, frontalAreaM2: c.frontalAreaM2,
      });
      const slice = addon.readSliceRGBA(this.ctx, AXIS_CODE[c.sliceAxis] || 0, c.sliceFraction,
        FIELD_CODE[c.fieldMode] === undefined ? 1 : FIELD_CODE[c.fieldMode], c.sliceRules === 'cpu' ? 0 : 1,
        c.uInletLattice);
      const out = { width: slice.width, height: slice.height, rgba: slice.rgba, stepCount: this.stepCount,
        forceReadout, forces: f };
      if (readMacroField) out.macroField = addon.readMacro(this.ctx, this.voxel.nx * this.voxel.ny * this.voxel.nz);
      return out;
    } finally {
      this.isBusy = false;
    }
  }

  /** One frame of the reference's solver loop (main.ts:491-512): stepAndExtractSlice over `substepsPerFrame`
   * substeps (4 with a GPU), the frame time dtSec (at least 0.1 ms) and MLUPS = cells * substeps / (dtSec * 1e6). */
  async tick(needMacro = false, substepsPerFrame = 4) {
    const t0 = process.hrtime.bigint();
    const out = await this.stepAndExtractSlice(substepsPerFrame, needMacro);
    if (!out) return null;
    const dtSec = Math.max(1e-4, Number(process.hrtime.bigint() - t0) / 1e9);
    out.mlups = (this.voxel.nx * this.voxel.ny * this.voxel.nz * substepsPerFrame) / (dtSec * 1e6);
    out.frameSeconds = dtSec;
    this.latestForceReadout = out.forceReadout;
    return out;
  }

  /** computeQCriterion (vortexIso.ts:22-119) over the current field, on the GPU. */
  computeQCriterion() {
    if (!this.ctx || !this.voxel) return null;
    return addon.qCriterion(this.ctx, this.voxel.nx * this.voxel.ny * this.voxel.nz, this.voxel.dx);
  }

  /** Point probe readouts (pointProbe.ts:284-355) at world points [{x, y, z}], with optional per-point
   * mesh-surface hits. */
  probe(points, onSurface = null) {
    if (!this.ctx || !this.voxel) return null;
    const xyz = new Float64Array(points.length * 3);
    points.forEach((p, i) => { xyz[3 * i] = p.x; xyz[3 * i + 1] = p.y; xyz[3 * i + 2] = p.z; });
    const v = this.voxel, c = this.config;
    const raw = addon.probe(this.ctx, xyz, onSurface ? Uint8Array.from(onSurface, (b) => (b ? 1 : 0)) : null, {
      originX: v.originX, originY: v.originY, originZ: v.originZ, dx: v.dx, windSpeedMps: c.windSpeedMps,
      uInletLattice: c.uInletLattice, fluidDensity: c.fluid.density,
    });
    return points.map((p, i) => {
      const r = raw.subarray(13 * i, 13 * i + 13);
      return {
        worldPos: { x: p.x, y: p.y, z: p.z }, latticeCell: { ix: r[0], iy: r[1], iz: r[2] },
        speedMps: r[3], speedMph: r[4], velocityMps: { ux: r[5], uy: r[6], uz: r[7] }, gaugePressurePa: r[8],
        pressureCoeffCp: r[9], densityRho: r[10], dynamicPressurePa: r[11], isInsideSolid: r[12] === 1,
      };
    });
  }

  destroy() {
    if (this.ctx) addon.destroy(this.ctx);
    this.ctx = null;
  }
}

/**
 * parseSTL (stlLoader.ts:9-38): binary or ASCII STL bytes (ArrayBuffer, Buffer or typed array) to a
 * geometry `{ position: Float32Array }` of unindexed triangles. Throws on bytes the reference rejects.
 */
function parseSTL(buffer) {
  return { position: addon.parseSTL(buffer) };
}

/**
 * normalizeAndCenterGeometry (stlLoader.ts:101-127): centres the geometry's bounding box on the origin and
 * rescales mm-sized parts, in place. Returns `{ geometry, autoScaledFromMm }`.
 */
function normalizeAndCenterGeometry(geometry) {
  const r = addon.normalizeAndCenterGeometry(geometry.position);
  return { geometry, autoScaledFromMm: r.autoScaledFromMm };
}

module.exports = {
  LBMEngine, ForceConvergenceTracker, voxelizeMeshToLattice, parseSTL, normalizeAndCenterGeometry,
  exportTelemetryToCSV, exportRunReportToJSON, computeAerodynamicTelemetry, addon,
};
