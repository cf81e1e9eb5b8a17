"""Reader for HipFT's ``hipft.in`` namelist (&hipft_input_parameters).

The reference reads one Fortran namelist group (src/hipft.f90:8324-8406,
defaults in module input_parameters 493-726, documented in
doc/hipft.in.documentation).  This reader handles the subset of namelist
syntax those files use: ``name = value`` pairs, comma-separated lists,
quoted strings, ``.true.``/``.false.``, Fortran ``d``/``e`` exponents and
``!`` comments.  Keys are returned lower-case; lists stay lists.
"""
import re

_TOKEN = re.compile(r"""'[^']*'|"[^"]*"|[^,\s]+""")


def _value(tok):
    t = tok.strip()
    if t[:1] in "'\"":
        return t[1:-1]
    low = t.lower()
    if low in (".true.", ".t.", "t", ".true"):
        return True
    if low in (".false.", ".f.", "f", ".false"):
        return False
    try:
        if re.fullmatch(r"[+-]?\d+", t):
            return int(t)
        return float(low.replace("d", "e"))
    except ValueError:
        return t


def _strip_comment(line):
    out, q = [], None
    for ch in line:
        if q:
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
        elif ch == "!":
            break
        out.append(ch)
    return "".join(out)


def parse(text, group="hipft_input_parameters"):
    """Parse the namelist group from text into a dict."""
    body = []
    inside = False
    for raw in text.splitlines():
        line = _strip_comment(raw).strip()
        if not line:
            continue
        if not inside:
            if line.lower().startswith("&" + group):
                inside = True
                line = line[len(group) + 1:].strip()
            else:
                continue
        if line.startswith("/") or line.lower().startswith("&end"):
            break
        body.append(line)
    joined = " ".join(body)
    out = {}
    # split on `name =` boundaries
    parts = re.split(r"([A-Za-z_][A-Za-z0-9_]*(?:\(\s*\d+\s*\))?)\s*=", joined)
    for i in range(1, len(parts) - 1, 2):
        key = parts[i].strip().lower()
        toks = [t for t in _TOKEN.findall(parts[i + 1]) if t != ","]
        vals = [_value(t) for t in toks]
        if not vals:
            continue
        out[key] = vals[0] if len(vals) == 1 else vals
    return out


def read(path):
    with open(path) as fh:
        return parse(fh.read())


# Namelist array names whose per-realization lists the driver takes as-is.
LIST_KEYS = {
    "diffusion_coef_constants", "flow_dr_coef_p0_values", "flow_dr_coef_p2_values",
    "flow_dr_coef_p4_values", "flow_mf_coef_p1_values", "flow_mf_coef_p3_values",
    "flow_mf_coef_p5_values", "flow_attenuate_values", "assimilate_data_mu_powers",
    "assimilate_data_mu_limits", "assimilate_data_lat_limits",
    "source_rfe_total_unsigned_flux_per_hours",
}


def to_run_config(nml, base_dir="."):
    """Map a parsed namelist onto driver.run's configuration dict.  An
    ``initial_map_filename`` is read (HDF5, hipft_amd.h5io) into the initial
    map and its scales, as load_initial_condition does (src/hipft.f90:1820-1870);
    relative paths are taken from base_dir."""
    import os
    cfg = {}
    for k, v in nml.items():
        if k in LIST_KEYS and not isinstance(v, list):
            v = [v]
        cfg[k] = v
    cfg["base_dir"] = base_dir  # flow and assimilation lists are read relative to it
    fname = cfg.get("initial_map_filename")
    if fname:
        from . import h5io
        path = fname if os.path.isabs(fname) else os.path.join(base_dir, fname)
        data, scales = h5io.read_map(path)
        cfg["initial_map"] = data
        cfg["initial_scales"] = scales[:2]
    return cfg
