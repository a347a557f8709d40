"""ASCII light-curve files: Lcurve::Data::rasc / wrasc (lcurve.cpp:14-100).

One point per line, ``time expose flux ferr weight ndiv``.  Lines that
start with '#', a space, a tab or are empty are skipped.  A negative
uncertainty marks a point to be ignored: ferr becomes 0 and weight 0
(lcurve.cpp:26-29).  Writing uses the reference's fixed-point layout
(lcurve.cpp:35-44): time with 8 decimals in 17 columns, expose with 8,
flux with 5 decimals in 10 columns, ferr and weight with 5, ndiv.
"""
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
