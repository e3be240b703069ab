"""Writer for ``pds_input.txt`` decks from a plain-dict problem spec.

The keyword/value-line format and every keyword name follow the reference
parser (pdDataManager.cpp:4-758, pdDataManager.h:270-421): a keyword on its
own line, the values on the next line, whitespace separated.  Keywords the
reference does not know (``grid_jitter``, ``number_of_slits``/``slit_N``,
``traction_as_body_force``) are extensions; the reference's search-by-keyword
parser skips lines it is not asked for, so a deck stays readable by it.

Spec keys (all lengths in metres, moduli in Pa):
  title, grid_dimension (nx,ny,nz), grid_spacing, grid_center (3),
  horizon, total_time, total_step, step_dump_frequency, safety_factor,
  fixed_time_step (None = stable dt),
  materials: [{type: 1, E, sigy, ecrit, rho, gravity: (3)}],
  bcs: [{type, flags: (3), value: (3 or 9), t_end, load}],
  material_spaces / deletion_spaces / boundary_spaces:
      [{geo: 1, ref, box: (x1lo,x1hi,x2lo,x2hi,x3lo,x3hi)} |
       {geo: 2, ref, cyl: (r, cx, cy, x3lo, x3hi)}]   (ref is 1-based, as in the deck)
  damping, bdryeff_on, cci: (on, modified) | None, fgi: (on, intg_point) | None,
  ai: (on, mls_on, intg_point) | None (used when cci and fgi are off),
  support_domain_size (None = 3 dx),
  jitter (fraction of dx), seed, slits: [{type: 1, box: (6)} | {type: 2, c, n, t, half_t, half_s}],
  traction_body (0/1)
"""

from __future__ import annotations

import math
from typing import Any, Dict, List

import numpy as np


def _fmt(v: Any) -> str:
    if isinstance(v, bool):
        return "1" if v else "0"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, str):
        return v
    return repr(float(v))


def _line(*vals: Any) -> str:
    return " " + " ".join(_fmt(v) for v in vals)


def deck_text(spec: Dict[str, Any], crlf: bool = False) -> str:
    """Render a spec dict as pds_input.txt text (LF by default, CRLF on request)."""
    out: List[str] = []

    def kw(name: str, *vals: Any) -> None:
        out.append(name)
        out.append(_line(*vals))

    out.append("// generated deck")
    kw("problem_title", spec.get("title", "pdm_problem"))
    kw("grid_center", *spec.get("grid_center", (0.0, 0.0, 0.0)))
    kw("grid_dimension", *[int(v) for v in spec["grid_dimension"]])
    kw("grid_spacing", spec["grid_spacing"])
    if spec.get("horizon") is not None:
        kw("horizon", spec["horizon"])
    kw("total_time", spec["total_time"])
    kw("total_step", int(spec["total_step"]))
    kw("step_dump_frequency", int(spec.get("step_dump_frequency", 1000000)))
    if spec.get("safety_factor") is not None:
        kw("safety_factor", spec["safety_factor"])
    if spec.get("fixed_time_step") is not None:
        kw("fixed_time_step", spec["fixed_time_step"])
    mats = spec["materials"]
    kw("number_of_materials", len(mats))
    for i, m in enumerate(mats, 1):
        kw(f"material_{i}", int(m.get("type", 1)), m["E"], m.get("sigy", 1e30), m["ecrit"], m["rho"])
        g = m.get("gravity")
        if g is not None:
            kw(f"gravity_{i}", *g)
    bcs = spec.get("bcs", [])
    kw("number_of_boundary_conditions", len(bcs))
    for i, b in enumerate(bcs, 1):
        vals = list(b["value"])
        kw(f"boundary_condition_{i}", int(b["type"]), *[int(f) for f in b["flags"]], *vals, b["t_end"], b["load"])

    def spaces(prefix: str, count_kw: str, lst: List[Dict[str, Any]], with_ref: bool) -> None:
        kw(count_kw, len(lst))
        for i, s in enumerate(lst, 1):
            geo = int(s.get("geo", 1))
            ref = [int(s["ref"])] if with_ref else []
            body = s["box"] if geo == 1 else s["cyl"]
            kw(f"{prefix}{i}", geo, *ref, *body)

    spaces("material_space_", "number_of_material_spaces", spec["material_spaces"], True)
    if spec.get("deletion_spaces"):
        spaces("deletion_space_", "number_of_deletion_spaces", spec["deletion_spaces"], False)
    spaces("boundary_space_", "number_of_boundary_spaces", spec.get("boundary_spaces", []), True)
    kw("damping_coeff", spec.get("damping", 0.0))
    kw("bdryEff_on", int(spec.get("bdryeff_on", 1)))
    cci = spec.get("cci", (1, 1))
    if cci is not None and cci[0]:
        kw("cci_on", 1, int(cci[1]))
    else:
        kw("cci_on", 0, 0)
        fgi = spec.get("fgi")
        ai = spec.get("ai")
        if fgi is not None and fgi[0]:
            kw("fgi_on", 1)
            kw("intg_point", int(fgi[1]))
        elif ai is not None and ai[0]:
            kw("fgi_on", 0)
            kw("ai_on", 1, int(ai[1]))
            if len(ai) > 2 and ai[2] is not None:
                kw("intg_point", int(ai[2]))
        else:
            raise ValueError("spec needs cci, fgi or ai")
        if spec.get("support_domain_size") is not None:
            kw("support_domain_size", spec["support_domain_size"])
    if spec.get("jitter"):
        kw("grid_jitter", spec["jitter"], int(spec.get("seed", 0)))
    slits = spec.get("slits", [])
    if slits:
        kw("number_of_slits", len(slits))
        for i, s in enumerate(slits, 1):
            if int(s["type"]) == 1:
                kw(f"slit_{i}", 1, *s["box"])
            else:
                kw(f"slit_{i}", 2, *s["c"], *s["n"], *s["t"], s["half_t"], s["half_s"])
    if spec.get("traction_body"):
        kw("traction_as_body_force", 1)
    text = "\n".join(out) + "\n"
    if crlf:
        text = text.replace("\n", "\r\n")
    return text


def bar_spec(nx: int, ny: int, nz: int, dx: float = 1e-3, delta_over_dx: float = 3.015,
             E: float = 72e9, rho: float = 2440.0, steps: int = 200, safety: float = 0.8,
             traction: float = 1e6, jitter: float = 0.0, seed: int = 0, slits=None,
             cci_modified: int = 1, bdryeff_on: int = 1) -> Dict[str, Any]:
    """Uniaxial bar: x=0 face fixed (DISPL, step), +x traction on the x=L layer.

    The traction is a TRACTION bc (type 4) whose value is the total force
    sigma * (ny dx) * (nz dx); it acts only with ``traction_body`` on (the
    reference treats type 4 as a no-op).  Node centres sit at (i+0.5) dx.
    """
    L = nx * dx
    area = (ny * dx) * (nz * dx)
    big = 1e3
    spec: Dict[str, Any] = dict(
        title=f"bar_{nx}x{ny}x{nz}",
        grid_dimension=(nx, ny, nz),
        grid_spacing=dx,
        grid_center=(0.5 * nx * dx, 0.5 * ny * dx, 0.5 * nz * dx),
        horizon=delta_over_dx * dx,
        total_time=1.0,
        total_step=steps,
        step_dump_frequency=10 ** 9,
        safety_factor=safety,
        materials=[dict(type=1, E=E, sigy=1e30, ecrit=1e10, rho=rho)],
        bcs=[dict(type=1, flags=(1, 1, 1), value=(0.0, 0.0, 0.0), t_end=1.0, load=2),
             dict(type=4, flags=(1, 0, 0), value=(traction * area, 0.0, 0.0), t_end=1.0, load=2)],
        material_spaces=[dict(geo=1, ref=1, box=(-big, big, -big, big, -big, big))],
        boundary_spaces=[dict(geo=1, ref=1, box=(-big, 1.0 * dx, -big, big, -big, big)),
                         dict(geo=1, ref=2, box=(L - 1.0 * dx, big, -big, big, -big, big))],
        damping=0.0,
        bdryeff_on=bdryeff_on,
        cci=(1, cci_modified),
        jitter=jitter,
        seed=seed,
        slits=list(slits or []),
        traction_body=1,
    )
    return spec


def mid_slit(spec: Dict[str, Any], angle_deg: float, axis_t=(0.0, 0.0, 1.0), edge: bool = False) -> Dict[str, Any]:
    """A planar slit through the bar centre, normal in the x-y plane at angle_deg to x.

    edge=False: the slit crosses the whole section (the bar is cut in two).
    edge=True: an edge crack, from the centre line out through the y = min face
    (half of the section), so the specimen stays connected.
    """
    nx, ny, nz = spec["grid_dimension"]
    dx = spec["grid_spacing"]
    a = math.radians(angle_deg)
    c = list(spec["grid_center"])
    big = 10.0 * max(nx, ny, nz) * dx
    n = (math.cos(a), math.sin(a), 0.0)
    if edge:
        # in-plane axis w = n x t = (sin a, -cos a, 0) for t = z: points to the y = min side
        w = (n[1] * axis_t[2] - n[2] * axis_t[1], n[2] * axis_t[0] - n[0] * axis_t[2], n[0] * axis_t[1] - n[1] * axis_t[0])
        if w[1] > 0:
            w = tuple(-x for x in w)
        c = [c[k] + big * w[k] for k in range(3)]
    return dict(type=2, c=tuple(c), n=n, t=tuple(axis_t), half_t=big, half_s=big)


def specimen_specs(count: int, seed: int = 0, nodes=(1e3, 1e5), steps: int = 1000, jitter: float = 0.01,
                   min_side: int = 5, safety: float = 0.8):
    """Random specimens for batched runs, each a bar spec with an edge crack.

    Per specimen, from numpy's default_rng(seed): node count log-uniform in
    ``nodes``; two box aspect ratios (length : width, width : depth) uniform in
    [1, 5]; delta / dx from {3.015, 4.015}; crack normal at an angle uniform in
    [0, 90] degrees to x; E uniform in [50, 200] GPa; node jitter ``jitter`` dx
    with the specimen's own seed.  The x = 0 layer is fixed and the x = L layer
    pulled with 1 MPa, as in ``bar_spec``.  Every side holds at least
    ``min_side`` nodes: thinner sections (2 or 3 layers, all of them surface
    layers) take bdryEff factors that make the stable-dt estimate of the
    reference unstable, so such boxes are thickened (the node count then exceeds
    the drawn one).
    """
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = math.exp(rng.uniform(math.log(nodes[0]), math.log(nodes[1])))
        r1, r2 = rng.uniform(1.0, 5.0, size=2)
        nz = max(min_side, int(round((n / (r1 * r2 * r2)) ** (1.0 / 3.0))))
        ny = max(min_side, int(round(r2 * nz)))
        nx = max(min_side, int(round(r1 * ny)))
        delta = float(rng.choice([3.015, 4.015]))
        angle = float(rng.uniform(0.0, 90.0))
        E = float(rng.uniform(50e9, 200e9))
        spec = bar_spec(nx, ny, nz, E=E, delta_over_dx=delta, steps=steps, jitter=jitter, safety=safety,
                        seed=int(rng.integers(0, 2 ** 31)))
        spec["title"] = f"specimen_{k}"
        spec["slits"] = [mid_slit(spec, angle, edge=True)]
        out.append(spec)
    return out


def plate_spec(nx: int = 400, ny: int = 400, nz: int = 50, jitter: float = 0.01, seed: int = 0, steps: int = 500,
               crack_fraction: float = 0.5) -> Dict[str, Any]:
    """Plate with two crossing slits through the thickness at its centre.

    Normals (1, 0, 0) and (0, 1, 0); each slit spans ``crack_fraction`` of the plate's
    width across its line (centred), so the plate stays in one piece.  Boundary
    conditions as in ``bar_spec``: x = 0 layer fixed, 1 MPa on the x = L layer.
    """
    spec = bar_spec(nx, ny, nz, jitter=jitter, seed=seed, steps=steps)
    spec["title"] = f"plate_{nx}x{ny}x{nz}"
    dx = spec["grid_spacing"]
    c = tuple(spec["grid_center"])
    big = 10.0 * max(nx, ny, nz) * dx
    spec["slits"] = [
        dict(type=2, c=c, n=(1.0, 0.0, 0.0), t=(0.0, 0.0, 1.0), half_t=big, half_s=0.5 * crack_fraction * ny * dx),
        dict(type=2, c=c, n=(0.0, 1.0, 0.0), t=(0.0, 0.0, 1.0), half_t=big, half_s=0.5 * crack_fraction * nx * dx),
    ]
    return spec
