"""Run telemetry export in the reference's formats: the RFC 4180 CSV of exportTelemetryToCSV and the JSON
report of exportRunReportToJSON (src/solver/domainWorkflows.ts:323-402), plus the aerodynamic telemetry that
fills a report's metadata (computeAerodynamicTelemetry, src/solver/fluidProperties.ts:7-52).

The reference formats numbers with JavaScript: Number.prototype.toFixed (the exact binary value rounded to
the nearest multiple of 10^-d, ties away from zero, "-" kept for negative values that round to zero but not
for -0), Number::toString (shortest round-trip digits, plain notation for exponents -7 < e < 21) and
JSON.stringify (non-finite numbers as null, undefined members left out, two-space indentation). The helpers
below follow those rules so that files written here read the same as the reference's.
"""
import datetime
import json
import math
from decimal import ROUND_HALF_UP, Decimal

NEWTONS_TO_LBF = 0.224808943   # lbmConstants.ts:101
NEWTONS_TO_KGF = 0.101971621   # lbmConstants.ts:102


def js_to_fixed(x, digits):
    """Number.prototype.toFixed(digits) for 0 <= digits <= 100."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if abs(x) >= 1e21 or math.isinf(x):
        return js_number_str(x)
    neg = x < 0
    q = Decimal(abs(x)).quantize(Decimal(1).scaleb(-digits), rounding=ROUND_HALF_UP)
    s = f"{q:f}"
    return ("-" if neg else "") + s


def js_number_str(x):
    """Number::toString(10) of a double."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    if x == 0:
        return "0"
    sign = "-" if x < 0 else ""
    d = Decimal(repr(abs(x))).normalize()  # repr: the shortest digits that round-trip
    t = d.as_tuple()
    digits = "".join(str(v) for v in t.digits)
    k = len(digits)
    n = k + t.exponent  # value = 0.digits * 10^n
    if k <= n <= 21:
        return sign + digits + "0" * (n - k)
    if 0 < n <= 21:
        return sign + digits[:n] + "." + digits[n:]
    if -6 < n <= 0:
        return sign + "0." + "0" * (-n) + digits
    e = n - 1
    es = ("+" if e >= 0 else "-") + str(abs(e))
    if k == 1:
        return sign + digits + "e" + es
    return sign + digits[0] + "." + digits[1:] + "e" + es


UNDEFINED = object()  # a member JSON.stringify leaves out


def _json(v, indent, level):
    pad, inner = " " * (indent * level), " " * (indent * (level + 1))
    if v is None:
        return "null"
    if v is True:
        return "true"
    if v is False:
        return "false"
    if isinstance(v, (int, float)):
        return js_number_str(v) if math.isfinite(float(v)) else "null"
    if isinstance(v, str):
        return json.dumps(v, ensure_ascii=False)
    if isinstance(v, dict):
        items = [(k, x) for k, x in v.items() if x is not UNDEFINED]
        if not items:
            return "{}"
        body = ",\n".join(f"{inner}{json.dumps(str(k), ensure_ascii=False)}: {_json(x, indent, level + 1)}"
                          for k, x in items)
        return "{\n" + body + "\n" + pad + "}"
    if isinstance(v, (list, tuple)):
        if not v:
            return "[]"
        body = ",\n".join(inner + ("null" if x is UNDEFINED else _json(x, indent, level + 1)) for x in v)
        return "[\n" + body + "\n" + pad + "]"
    raise TypeError(f"cannot serialise {type(v).__name__}")


def js_json_stringify(v, indent=2):
    """JSON.stringify(v, null, indent) for dicts, lists, str, numbers, bools and None (null)."""
    return _json(v, indent, 0)


def iso_now():
    """new Date().toISOString()"""
    t = datetime.datetime.now(datetime.timezone.utc)
    return t.strftime("%Y-%m-%dT%H:%M:%S.") + f"{t.microsecond // 1000:03d}Z"


def _get(d, *path):
    """Optional chaining: the value at path, or None when a link is missing or None."""
    for k in path:
        if d is None:
            return None
        d = d.get(k) if isinstance(d, dict) else getattr(d, k, None)
    return d


def _first(*vals):
    """a ?? b ?? c"""
    for v in vals:
        if v is not None:
            return v
    return None


def export_telemetry_to_csv(history, metadata):
    """exportTelemetryToCSV (domainWorkflows.ts:323-371): comment header, column header and one row per
    sample, CRLF line ends. history: dicts with step, timeSec, cd, cl and optional cm, dragForceN, liftForceN,
    xcpMeters; metadata: dict of SimulationRunMetadata fields (any may be missing)."""
    m = metadata or {}

    def fx(path, d, dflt):
        v = _get(m, *path)
        return js_to_fixed(v, d) if v is not None else dflt

    lines = ["# AeroSim Aerodynamic Wind Tunnel Telemetry Export",
             f"# Generated: {_first(m.get('timestamp'), iso_now())}",
             f"# Specimen: {_first(m.get('specimenName'), 'Specimen')}",
             f"# Wind Speed: {fx(('windSpeedMps',), 1, '0.0')} m/s",
             f"# Reynolds Number: {fx(('reynolds',), 0, '0')}",
             f"# Frontal Area: {fx(('meshStats', 'frontalArea'), 4, '0')} m2",
             f"# Mean Cd: {fx(('meanCd',), 4, '0')}",
             f"# Mean Cl: {fx(('meanCl',), 4, '0')}",
             f"# Pitching Moment Cm: {fx(('meanCm',), 4, '0')}",
             f"# Center of Pressure X: {fx(('centerOfPressureX',), 3, '0')} m",
             "",
             "step,time_sec,cd,cl,cm,drag_force_n,drag_force_lbf,drag_force_kgf,lift_force_n,lift_force_lbf,xcp_meters"]
    for s in history:
        drag = _first(s.get("dragForceN"), m.get("dragForceN"), 0.0)
        lift = _first(s.get("liftForceN"), m.get("liftForceN"), 0.0)
        row = [js_number_str(s["step"]), js_to_fixed(s["timeSec"], 4), js_to_fixed(s["cd"], 5), js_to_fixed(s["cl"], 5),
               js_to_fixed(_first(s.get("cm"), m.get("meanCm"), 0.0), 5), js_to_fixed(drag, 2),
               js_to_fixed(drag * NEWTONS_TO_LBF, 2), js_to_fixed(drag * NEWTONS_TO_KGF, 2), js_to_fixed(lift, 2),
               js_to_fixed(lift * NEWTONS_TO_LBF, 2),
               js_to_fixed(_first(s.get("xcpMeters"), m.get("centerOfPressureX"), 0.0), 3)]
        lines.append(",".join(row))
    return "\r\n".join(lines)


def _num_fixed(v, d):
    """Number(v.toFixed(d))"""
    return float(js_to_fixed(v, d))


REPORT_APPLICATION = "AeroSim \u2014 3D WebGPU Wind Tunnel Platform"  # the report schema's producer field
REPORT_VERSION = "0.1.0"


def export_run_report_to_json(metadata, history, export_timestamp=None, application=REPORT_APPLICATION,
                              version=REPORT_VERSION):
    """exportRunReportToJSON (domainWorkflows.ts:376-402): the report object with the run's metadata, units and
    rounded telemetry history, as JSON.stringify(payload, null, 2). The producer name and version default to the
    reference report's, which its consumers and tests key on."""
    hist = []
    for s in history:
        hist.append({
            "step": s["step"],
            "timeSec": _num_fixed(s["timeSec"], 4),
            "cd": _num_fixed(s["cd"], 5),
            "cl": _num_fixed(s["cl"], 5),
            "cm": _num_fixed(s["cm"], 5) if s.get("cm") is not None else UNDEFINED,
            "dragForceN": _num_fixed(s["dragForceN"], 2) if s.get("dragForceN") is not None else UNDEFINED,
            "liftForceN": _num_fixed(s["liftForceN"], 2) if s.get("liftForceN") is not None else UNDEFINED,
            "xcpMeters": _num_fixed(s["xcpMeters"], 3) if s.get("xcpMeters") is not None else UNDEFINED,
        })
    run = dict(metadata)
    run["units"] = {"windSpeed": "m/s", "forces": "Newtons", "pressure": "Pascals", "lengths": "meters",
                    "areas": "square_meters"}
    run["telemetryHistory"] = hist
    payload = {"application": application, "version": version,
               "exportTimestamp": export_timestamp or iso_now(), "simulationRun": run}
    return js_json_stringify(payload, 2)


def compute_aerodynamic_telemetry(wind_speed, characteristic_length, frontal_area, fluid, tunnel):
    """computeAerodynamicTelemetry (fluidProperties.ts:7-52). fluid: dict with density, kinematicViscosity,
    speedOfSound; tunnel: dict or Tunnel with width and height."""
    speed = max(0.0, wind_speed)
    length = max(1e-6, characteristic_length)
    nu = max(1e-12, fluid["kinematicViscosity"])
    sound = max(1.0, fluid["speedOfSound"])
    re = (speed * length) / nu
    width, height = _get(tunnel, "width"), _get(tunnel, "height")
    cross = max(1e-4, width * height)
    regime = "Laminar"
    if re >= 3.5e5:
        regime = "Turbulent (Supercritical)"
    elif re >= 2000:
        regime = "Turbulent (Subcritical)"
    elif re >= 500:
        regime = "Transitional"
    return {"reynoldsNumber": re, "machNumber": speed / sound, "dynamicPressurePa": 0.5 * fluid["density"] * speed * speed,
            "frontalAreaM2": frontal_area, "characteristicLengthM": characteristic_length,
            "blockageRatioPercent": (frontal_area / cross) * 100, "flowRegime": regime}


def history_for_export(readout, wind_speed, fluid_density, frontal_area):
    """The samples main.ts:306-315 hands to the exporters: timeSec = step * 0.001, forces from the tracker's
    history (or Cd and Cl times q * A), cm and xcp from the sample or the readout's means."""
    q_area = 0.5 * fluid_density * wind_speed * wind_speed * frontal_area
    out = []
    for h in readout["history"]:
        out.append({"step": h["step"], "timeSec": h["step"] * 0.001, "cd": h["cd"], "cl": h["cl"],
                    "cm": _first(h.get("cm"), readout["pitchingMomentCm"]),
                    "dragForceN": _first(h.get("dragForceN"), h["cd"] * q_area),
                    "liftForceN": _first(h.get("liftForceN"), h["cl"] * q_area),
                    "xcpMeters": _first(h.get("xcp"), readout["centerOfPressureX"])})
    return out
