"""ASCII light-curve files: Lcurve::Data::rasc / wrasc (lcurve.cpp:14-100).

One point per line, ``time expose flux ferr weight ndiv``.  Lines that
start with '#', a space, a tab or are empty are skipped.  A negative
uncertainty marks a point to be ignored: ferr becomes 0 and weight 0
(lcurve.cpp:26-29).  Writing uses the reference's fixed-point layout
(lcurve.cpp:35-44): time with 8 decimals in 17 columns, expose with 8,
flux with 5 decimals in 10 columns, ferr and weight with 5, ndiv.
"""
import json
import math
import sys

import numpy as np

from . import make_data, data_view


def read_data(path):
    """Returns an lc_datum array (see make_data) read from an ASCII file."""
    rows = []
    with open(path, "r") as f:
        for n, line in enumerate(f, 1):
            if line == "" or line[0] in "# \t\n":
                continue
            parts = line.split()
            try:
                t, e, fl, fe, w = (float(x) for x in parts[:5])
                nd = int(parts[5])
            except (ValueError, IndexError):
                raise RuntimeError("Lcurve::operator>>: failed to read t,e,f,fe,wgt,ndiv of a Datum "
                                   f"(line {n} of {path})") from None
            if fe < 0.0:
                fe, w = 0.0, 0.0
            rows.append((t, e, fl, fe, w, nd))
    if not rows:
        return make_data(np.zeros(0))
    a = np.array(rows, dtype=np.float64)
    return make_data(a[:, 0], expose=a[:, 1], flux=a[:, 2], ferr=a[:, 3], weight=a[:, 4],
                     ndiv=a[:, 5].astype(np.int32))


def write_data(path, data):
    """Writes an lc_datum array in the reference's ASCII layout."""
    v = data_view(data)
    with open(path, "w") as f:
        for r in v:
            f.write(f"{r['time']:17.8f} {r['expose']:.8f} {r['flux']:10.5f} {r['ferr']:.5f} "
                    f"{r['weight']:.5f} {int(r['ndiv'])}\n")


# Parameters written back into a configuration, in the reference's order
# (new_helpers.h:480-533, write_config_and_model_to_json).
CONFIG_PARAMS = (
    "q iangle r1 r2 cphi3 cphi4 spin1 spin2 t1 t2 "
    "ldc1_1 ldc1_2 ldc1_3 ldc1_4 ldc2_1 ldc2_2 ldc2_3 ldc2_4 "
    "velocity_scale beam_factor1 beam_factor2 t0 period pdot deltat "
    "gravity_dark1 gravity_dark2 absorb slope quad cube third "
    "rdisc1 rdisc2 height_disc beta_disc temp_disc texp_disc lin_limb_disc quad_limb_disc temp_edge absorb_edge "
    "radius_spot length_spot height_spot expon_spot epow_spot angle_spot yaw_spot temp_spot tilt_spot cfrac_spot "
    "stsp11_long stsp11_lat stsp11_fwhm stsp11_tcen stsp12_long stsp12_lat stsp12_fwhm stsp12_tcen "
    "stsp13_long stsp13_lat stsp13_fwhm stsp13_tcen stsp21_long stsp21_lat stsp21_fwhm stsp21_tcen "
    "stsp22_long stsp22_lat stsp22_fwhm stsp22_tcen uesp_long1 uesp_long2 uesp_lathw uesp_taper uesp_temp").split()


def pparam_string(p):
    """'value range dstep vary defined' as new_helpers.h:23-30 writes it
    (value with 10 decimals, range and dstep with 6)."""
    return "%.10f %.6f %.6f %d %d" % (p.value, p.range, p.dstep, int(p.vary), int(p.defined))


def write_config_and_model(model, config, path):
    """A copy of `config` (dict or JSON text) with its model_parameters set
    from `model` (Helpers::write_config_and_model_to_json,
    new_helpers.h:470-541): pretty-printed with two-space indents and keys
    in sorted order, as nlohmann::json prints an object."""
    import copy
    import json
    cfg = json.loads(config) if isinstance(config, str) else copy.deepcopy(config)
    mp = cfg.get("model_parameters")
    if not isinstance(mp, dict):
        raise ValueError("write_config_and_model: config misses 'model_parameters' object")
    for name in CONFIG_PARAMS:
        mp[name] = pparam_string(getattr(model.pod, name))
    with open(path, "w") as f:
        f.write(json.dumps(cfg, indent=2, sort_keys=True, ensure_ascii=False) + "\n")


def write_chain(path, names, pars, chisq, log_prior=None, thin=1):
    """One MCMC chain as the reference's sampler writes it
    (mcmc_solver.cpp:1488-1506): header step,<names>,chisq,log_prior, then
    every thin-th recorded state with the stream's default formatting (six
    significant digits).  pars: [steps, d]; log_prior defaults to 0 (no
    physical prior)."""
    pars = np.asarray(pars, float)
    chisq = np.asarray(chisq, float)
    lp = np.zeros(len(pars)) if log_prior is None else np.asarray(log_prior, float)
    thin = max(1, int(thin))
    with open(path, "w") as f:
        f.write("step" + "".join("," + n for n in names) + ",chisq,log_prior\n")
        for i in range(0, len(pars), thin):
            f.write("%d%s,%s,%s\n" % (i, "".join("," + format(v, "g") for v in pars[i]), format(chisq[i], "g"),
                                      format(lp[i], "g")))


# ------------------------------------------------------------ plot frames
def normalized_plot_device(device):
    """Plot device name trimmed of whitespace and quotes, lower case
    (new_helpers.h:35-46)."""
    return str(device).strip(" \t\n\r\v\f\"'").lower()


def plotting_streamed(device):
    """True for the machine-readable stream backend (new_helpers.h:52-56)."""
    return normalized_plot_device(device) in ("stream", "json-stream", "astra")


def plotting_disabled(device):
    """True when the device switches plotting off (new_helpers.h:62-69)."""
    return normalized_plot_device(device) in ("", "none", "null", "disable", "disabled", "off", "false", "no")


def _json_number(x):
    x = float(x)
    return x if math.isfinite(x) else None   # nlohmann dumps NaN and inf as null


def model_frame(data, fit, no_file=False, copy=None, metadata=None):
    """The JSON text of one "lcurve.plot.v1" frame (new_helpers.h:71-109):
    x (time minus an origin when the span is below 1% of the mean time),
    flux, error (ferr of `copy` unless no_file; non-positive or non-finite
    ones become 1e-3), model, residual, and meta when given; keys sorted and
    no spaces, as nlohmann::json::dump() writes them.  None when there is
    nothing to plot."""
    v = data_view(data)
    fit = np.asarray(fit, np.float64)
    n = len(v)
    if n == 0 or fit.size < n:
        return None
    t = v["time"]
    x1, x2 = float(np.min(t)), float(np.max(t))
    origin = x1 if x2 - x1 < 0.01 * abs((x1 + x2) / 2.0) else 0.0
    src = v if (no_file or copy is None) else data_view(copy)
    frame = {"protocol": "lcurve.plot.v1", "x": [], "flux": [], "error": [], "model": [], "residual": []}
    for i in range(n):
        sigma = float(src["ferr"][i])
        if not (sigma > 0.0) or not math.isfinite(sigma):
            sigma = 1e-3
        f = float(v["flux"][i])
        frame["x"].append(_json_number(float(t[i]) - origin))
        frame["flux"].append(_json_number(f))
        frame["error"].append(_json_number(sigma))
        frame["model"].append(_json_number(fit[i]))
        frame["residual"].append(_json_number((f - float(fit[i])) / sigma))
    if metadata:
        frame["meta"] = metadata
    return json.dumps(frame, sort_keys=True, separators=(",", ":"), allow_nan=False)


def stream_model_frame(data, fit, no_file=False, copy=None, metadata=None, out=None):
    """Writes "@@LCURVE_PLOT@@" + the frame + a newline to `out` (stdout)
    and flushes, as Helpers::stream_model_frame does for the stream plot
    device."""
    text = model_frame(data, fit, no_file, copy, metadata)
    if text is None:
        return
    out = sys.stdout if out is None else out
    out.write("@@LCURVE_PLOT@@" + text + "\n")
    out.flush()


def plot_model(data, fit, no_file, copy, device, out=None):
    """Helpers::plot_model (new_helpers.h:111-122) for the devices that do
    not open gnuplot: nothing when the device disables plotting, one frame
    with meta {"phase": "final"} for the stream device.  Returns True when a
    frame was written; gnuplot devices (an external program) return False
    without plotting."""
    if plotting_disabled(device):
        return False
    if plotting_streamed(device):
        stream_model_frame(data, fit, no_file, copy, {"phase": "final"}, out=out)
        return True
    return False
