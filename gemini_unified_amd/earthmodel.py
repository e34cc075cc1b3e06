"""Node earth model, external radial nodes and the (omega, l) grid.

Host-side set-up of the Green-FK path, mirroring the reference modules

* ``f90/nodeEarthmodel.f90`` (JSON node model, SI units; reader at :98-257),
* ``f90/externalRadialNodes.f90`` (node radii from blocks :82-156, splitting at
  discontinuities :337-418, closest-node lookup :538-547),
* ``f90/computeGreenFKSpectraForSynthetics.f90:186-236`` (frequency grid,
  ``sigma``, degree counts ``nwn``).

It also holds the seeded global and regional model generators that write the
reference JSON schema.  Arrays use the reference's ordering (knots bottom-up,
layers bottom-up) with 0-based Python indexing; integer *indices stored in the
model* (``iktop``) keep the reference's 1-based values so the C ABI receives
exactly what the Fortran arrays hold.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np

# FP32 literals used as FP64 values by the reference (mathConstants.f90:26-35)
MC_PI = float(np.float32(3.141592653589793))
MC_TWO_PID = 2.0 * MC_PI
DEPS = float(np.finfo(np.float64).eps)

ATTMODES = {"ELASTIC": 0, "ATTENUATION_ONLY": 1, "DISPERSION_ONLY": 2, "ATTENUATION_AND_DISPERSION": 3}


@dataclass
class NodeEarthmodel:
    """Mirror of ``type(node_earthmodel)`` (nodeEarthmodel.f90:48-83)."""

    rearth: float
    ric: float
    roc: float
    fref: float
    nic: int
    noc: int
    fsflag: int
    equidis: int
    aniflag: int
    rk: np.ndarray
    ro: np.ndarray
    vpv: np.ndarray
    vph: np.ndarray
    vsv: np.ndarray
    vsh: np.ndarray
    eta: np.ndarray
    qkinv: np.ndarray
    qminv: np.ndarray
    rb: np.ndarray
    iktop: np.ndarray      # 1-based knot index of each layer top
    iflso: np.ndarray
    description: str = ""

    @property
    def nk(self) -> int:
        return int(self.rk.size)

    @property
    def nlay(self) -> int:
        return int(self.rb.size)

    def layer_index(self, r: float) -> tuple[int, int]:
        """getLayerIndexNodeEarthmodel (nodeEarthmodel.f90:599-610): 1-based layer, top flag."""
        nl = locate(r, self.rb) + 1
        top = 1 if abs(r / self.rb[nl - 1] - 1.0) < DEPS else 0
        return nl, top

    # --- JSON (reference SI schema) ------------------------------------------------
    @classmethod
    def from_json(cls, path_or_dict) -> "NodeEarthmodel":
        """createNodeEarthmodel (nodeEarthmodel.f90:98-257)."""
        d = path_or_dict
        if not isinstance(d, dict):
            with open(d) as fh:
                d = json.load(fh)
        nlay = int(d["num_layers"])
        nk = int(d["num_nodes"])
        iflso = np.asarray(d["isfluid"], dtype=np.int32)
        fs = int(d["full_space"])
        if iflso[nlay - 1] == 1 and fs == 1:
            raise ValueError("Full space and ocean is not allowed!!")
        npr = d["node_properties"]
        rearth = float(d["earth_radius"])
        rk = rearth - np.asarray(npr["depth"], dtype=np.float64)
        qk = 1.0 / np.asarray(npr["Qkappa"], dtype=np.float64)
        qm = 1.0 / np.asarray(npr["Qmu"], dtype=np.float64)
        qm = np.where(qm < 0.0, 0.0, qm)
        ani = int(d["anisotropic"]) if not isinstance(d["anisotropic"], bool) else int(d["anisotropic"])
        if ani == 0:
            vp = np.asarray(npr["P-velocity"], dtype=np.float64)
            vs = np.asarray(npr["S-velocity"], dtype=np.float64)
            vpv, vph, vsv, vsh = vp, vp.copy(), vs, vs.copy()
            eta = np.ones(nk)
        else:
            vpv = np.asarray(npr["PV-velocity"], dtype=np.float64)
            vph = np.asarray(npr["PH-velocity"], dtype=np.float64)
            vsv = np.asarray(npr["SV-velocity"], dtype=np.float64)
            vsh = np.asarray(npr["SH-velocity"], dtype=np.float64)
            eta = np.asarray(npr["eta"], dtype=np.float64)
        return cls(
            rearth=rearth, ric=float(d["radius_inner_core"]), roc=float(d["radius_outer_core"]),
            fref=float(d["reference_frequency_hz"]), nic=int(d["layer_inner_core"]),
            noc=int(d["layer_outer_core"]), fsflag=fs, equidis=int(d["equidistant_nodes"]), aniflag=ani,
            rk=rk, ro=np.asarray(npr["density"], dtype=np.float64), vpv=vpv, vph=vph, vsv=vsv, vsh=vsh,
            eta=eta, qkinv=qk, qminv=qm,
            rb=np.asarray(d["radius_upper_layer_boundary_m"], dtype=np.float64),
            iktop=np.asarray(d["index_upper_boundary"], dtype=np.int32), iflso=iflso,
            description=str(d.get("description", "")),
        )


def locate(xx: float, x: np.ndarray) -> int:
    """locate_double_precision (locatePoint.f90:80-110); returns the 1-based j with
    x(j) < xx <= x(j+1) for ascending x (0 if xx <= x(1))."""
    n = len(x)
    lidx, ridx = 0, n + 1
    asc = x[n - 1] > x[0]
    while ridx - lidx > 1:
        midx = (lidx + ridx) // 2
        if asc:
            if xx > x[midx - 1]:
                lidx = midx
            else:
                ridx = midx
        else:
            if xx > x[midx - 1]:
                ridx = midx
            else:
                lidx = midx
    return lidx


# ----------------------------------------------------------------------------------
#  Seeded models
# ----------------------------------------------------------------------------------
def _layered_json(description, rearth, layers, fref=1.0, full_space=0, nic=0, noc=0, ric=0.0, roc=0.0):
    """Assemble the reference JSON schema from per-layer knot lists.

    ``layers`` is a list (bottom-up) of dicts with keys r (ascending radii),
    density, vp, vs, qk, qm (arrays of equal length) and fluid (bool).
    """
    depth, dens, vp, vs, qk, qm, lidx, iktop, rb, iflso = [], [], [], [], [], [], [], [], [], []
    for il, lay in enumerate(layers, start=1):
        depth += list(rearth - np.asarray(lay["r"], dtype=np.float64))
        dens += list(lay["density"]); vp += list(lay["vp"]); vs += list(lay["vs"])
        qk += list(lay["qk"]); qm += list(lay["qm"])
        lidx += [il] * len(lay["r"])
        iktop.append(len(depth))
        rb.append(float(lay["r"][-1]))
        iflso.append(1 if lay["fluid"] else 0)
    return {
        "description": description, "type": "nm", "reference_frequency_hz": fref, "anisotropic": 0,
        "earth_radius": rearth, "full_space": full_space, "equidistant_nodes": 1,
        "num_layers": len(layers), "num_nodes": len(depth), "layer_inner_core": nic,
        "layer_outer_core": noc, "radius_inner_core": ric, "radius_outer_core": roc,
        "radius_upper_layer_boundary_m": rb, "isfluid": iflso, "index_upper_boundary": iktop,
        "units": {"density": "kg/m3", "velocity": "m/s", "depth": "m"},
        "node_properties": {
            "layer_index": lidx, "depth": [float(v) for v in depth], "density": [float(v) for v in dens],
            "P-velocity": [float(v) for v in vp], "S-velocity": [float(v) for v in vs],
            "Qkappa": [float(v) for v in qk], "Qmu": [float(v) for v in qm],
        },
    }


def global_model_json(seed: int = 1, nk: int = 200) -> dict:
    """Seeded 1-D global model (R = 6371 km, ``nk`` knots).

    Layers bottom-up: central homogeneous sphere (1 knot, r <= 100 km), solid
    inner core to 1221.5 km, fluid outer core to 3480 km (vs = 0), solid mantle
    to the surface.  vp 5.8 -> 13.7 km/s, vs 3.2 -> 7.3 km/s, rho 2.6 -> 13.0
    g/cc, each piecewise linear in depth with seeded +-2 % knot perturbations;
    Qmu 80 - 600, Qkappa 57823, reference frequency 1 Hz.
    """
    rng = np.random.default_rng(seed)
    R, rsph, ric, roc = 6371000.0, 100000.0, 1221500.0, 3480000.0
    n_ic = max(4, int(round((nk - 1) * 0.10)))
    n_oc = max(4, int(round((nk - 1) * 0.25)))
    n_ma = nk - 1 - n_ic - n_oc

    def lin(r, r0, r1, v0, v1):
        return v0 + (v1 - v0) * (np.asarray(r) - r0) / (r1 - r0)

    def pert(v):
        return np.asarray(v) * (1.0 + 0.02 * rng.uniform(-1.0, 1.0, size=np.shape(v)))

    r_ic = np.linspace(rsph, ric, n_ic)
    r_oc = np.linspace(ric, roc, n_oc)
    r_ma = np.linspace(roc, R, n_ma)
    ic = dict(r=r_ic, fluid=False,
              vp=pert(lin(r_ic, 0, ric, 11.30e3, 11.00e3)), vs=pert(lin(r_ic, 0, ric, 3.70e3, 3.50e3)),
              density=pert(lin(r_ic, 0, ric, 13.00e3, 12.76e3)),
              qk=np.full(n_ic, 57823.0), qm=np.full(n_ic, 84.6))
    oc = dict(r=r_oc, fluid=True,
              vp=pert(lin(r_oc, ric, roc, 10.30e3, 8.00e3)), vs=np.zeros(n_oc),
              density=pert(lin(r_oc, ric, roc, 12.17e3, 9.90e3)),
              qk=np.full(n_oc, 57823.0), qm=np.full(n_oc, -1.0))
    ma = dict(r=r_ma, fluid=False,
              vp=pert(lin(r_ma, roc, R, 13.70e3, 5.80e3)), vs=pert(lin(r_ma, roc, R, 7.30e3, 3.20e3)),
              density=pert(lin(r_ma, roc, R, 5.57e3, 2.60e3)),
              qk=np.full(n_ma, 57823.0), qm=np.clip(lin(r_ma, roc, R, 80.0, 600.0), 80.0, 600.0))
    # keep the surface knot at the reference surface values (prem_1d_aski:5)
    ma["vp"][-1], ma["vs"][-1], ma["density"][-1] = 5800.0, 3200.0, 2600.0
    sph = dict(r=np.array([rsph]), fluid=False, vp=ic["vp"][:1].copy(), vs=ic["vs"][:1].copy(),
               density=ic["density"][:1].copy(), qk=np.array([57823.0]), qm=np.array([84.6]))
    return _layered_json(f"seeded global model, seed {seed}, {nk} knots", R, [sph, ic, oc, ma],
                         nic=2, noc=3, ric=ric, roc=roc)


def regional_model_json(seed: int = 6) -> dict:
    """Seeded regional model: 6 crust/upper-mantle layers on a homogeneous sphere (halfspace).

    Layer thicknesses are uniform in 2-40 km (scaled down if they add up to more than 200 km); the
    halfspace starts below the sixth layer.  vp 4.5 -> 8.2 km/s, vs 2.6 -> 4.6 km/s, rho 2.4 -> 3.4
    g/cc increase with depth over 0-200 km, linear within each layer (2 knots), with seeded +-2 %
    knot perturbations; Qmu 100 -> 300, Qkappa 57823.  The halfspace (layer 1) carries the values
    at the bottom of the sixth layer.
    """
    rng = np.random.default_rng(seed)
    R, zmax = 6371000.0, 200000.0
    th = rng.uniform(2000.0, 40000.0, size=6)
    if th.sum() > zmax:
        th *= zmax / th.sum()
    tops = np.concatenate([[0.0], np.cumsum(th)[:-1]])              # depths of the layer tops, surface first
    bots = np.cumsum(th)

    def prop(z, a, b):
        return a + (b - a) * (np.asarray(z) / zmax)

    def pert(v):
        return np.asarray(v) * (1.0 + 0.02 * rng.uniform(-1.0, 1.0, size=np.shape(v)))

    layers = []
    for zt, zb in reversed(list(zip(tops, bots))):                  # bottom-up
        z = np.array([zb, zt])
        layers.append(dict(r=R - z, fluid=False, vp=pert(prop(z, 4.5e3, 8.2e3)), vs=pert(prop(z, 2.6e3, 4.6e3)),
                           density=pert(prop(z, 2.4e3, 3.4e3)), qk=np.full(2, 57823.0),
                           qm=prop(z, 100.0, 300.0)))
    base = layers[0]
    sph = dict(r=np.array([R - bots[-1]]), fluid=False, vp=base["vp"][:1].copy(), vs=base["vs"][:1].copy(),
               density=base["density"][:1].copy(), qk=np.array([57823.0]), qm=base["qm"][:1].copy())
    return _layered_json(f"seeded regional model, seed {seed}", R, [sph] + layers, nic=0, noc=0)


def regional_model(seed: int = 6) -> NodeEarthmodel:
    return NodeEarthmodel.from_json(regional_model_json(seed))


def global_model(seed: int = 1, nk: int = 200) -> NodeEarthmodel:
    return NodeEarthmodel.from_json(global_model_json(seed, nk))


# ----------------------------------------------------------------------------------
#  External radial nodes (externalRadialNodes.f90)
# ----------------------------------------------------------------------------------
@dataclass
class ExternalRadialNodes:
    rearth: float
    rnod: np.ndarray                  # ascending radii (m)
    state: np.ndarray                 # 1 fluid, 0 solid
    layer: np.ndarray                 # 1-based layer index

    @property
    def nnod(self) -> int:
        return int(self.rnod.size)

    @classmethod
    def from_blocks(cls, nem: NodeEarthmodel, nnod_blocks, dr_blocks, shift=0.0, rearth=None):
        """createExternalRadialNodes (externalRadialNodes.f90:82-156)."""
        rearth = nem.rearth if rearth is None else rearth
        nnod = int(sum(nnod_blocks)) + 1
        prn = np.zeros(nnod)
        prn[nnod - 1] = rearth - shift
        n = nnod
        for nb, dr in zip(nnod_blocks, dr_blocks):
            for i in range(1, nb + 1):
                prn[n - i - 1] = prn[n - 1] - i * float(dr)
            n -= nb
        if np.any(prn < 0.0):
            raise ValueError("Negative radii, check your input")
        return cls.split_at_discos(nem, prn, rearth)

    @classmethod
    def split_at_discos(cls, nem: NodeEarthmodel, prn: np.ndarray, rearth: float):
        """splitNodesAtDiscosExternalRadialNodes (externalRadialNodes.f90:337-418)."""
        rb = nem.rb
        drmin = rb[-1] * 1.0e-4
        nnod = len(prn)
        rnod, lay, st = [], [], []
        ne = 0
        for ib in range(1, len(rb) + 1):
            n = locate(rb[ib - 1], prn)
            if n == 0:
                continue
            iflso = int(nem.iflso[ib - 1])
            rnod += list(prn[ne:n]); lay += [ib] * (n - ne); st += [iflso] * (n - ne)
            if n >= nnod:
                break
            up_close = abs(prn[n] - rb[ib - 1]) < drmin
            lo_close = abs(prn[n - 1] - rb[ib - 1]) < drmin
            if up_close and lo_close:
                raise ValueError("External nodes too close to each other, use bigger spacing")
            if up_close:
                rnod.append(rb[ib - 1] - drmin * 0.5); lay.append(ib); st.append(iflso)
                if n + 1 < nnod:
                    rnod.append(rb[ib - 1] + drmin * 0.5); lay.append(ib + 1); st.append(int(nem.iflso[ib]))
                    ne = n + 1
                else:
                    ne = n
            elif lo_close:
                rnod[-1] = rb[ib - 1] - drmin * 0.5; lay[-1] = ib; st[-1] = iflso
                rnod.append(rb[ib - 1] + drmin * 0.5); lay.append(ib + 1); st.append(int(nem.iflso[ib]))
                ne = n
            else:
                ne = n
                rnod.append(rb[ib - 1] - drmin * 0.5); lay.append(ib); st.append(iflso)
                rnod.append(rb[ib - 1] + drmin * 0.5); lay.append(ib + 1); st.append(int(nem.iflso[ib]))
        return cls(rearth=rearth, rnod=np.asarray(rnod, dtype=np.float64),
                   state=np.asarray(st, dtype=np.int32), layer=np.asarray(lay, dtype=np.int32))

    def node_index(self, rs: float) -> int:
        """getNodeIdxFromRadiusExternalRadialNodes (externalRadialNodes.f90:538-547), 1-based."""
        jr = locate(rs, self.rnod)
        if jr == 0:
            jr = 1
        if jr < self.nnod:
            if abs(rs - self.rnod[jr - 1]) > abs(rs - self.rnod[jr]):
                jr += 1
        return jr


# ----------------------------------------------------------------------------------
#  Frequency / degree grid (computeGreenFKSpectraForSynthetics.f90:186-236)
# ----------------------------------------------------------------------------------
@dataclass
class FrequencyGrid:
    df: float
    nf1: int
    nf2: int
    sigma: float
    nwn: np.ndarray = field(repr=False)      # degrees per frequency index jf = nf1..nf2 (l = 0..nwn-1)

    @property
    def jf(self) -> np.ndarray:
        return np.arange(self.nf1, self.nf2 + 1)

    @property
    def freqs(self) -> np.ndarray:
        return (self.jf - 1) * self.df

    @property
    def omre(self) -> np.ndarray:
        """omre = 2.d0*mc_pi*f with the FP32 pi constant."""
        return (2.0 * MC_PI) * self.freqs

    @property
    def nkf(self) -> int:
        return int(self.nwn.sum())

    @classmethod
    def global_fixed_lmax(cls, tlen: float, fmax: float, lmax: int, fmin: float = 0.0):
        df = 1.0 / tlen
        nf2 = int(np.rint(fmax / df)) + 1
        nf1 = max(int(np.rint(fmin / df)), 2)
        nwn = np.full(nf2 - nf1 + 1, lmax + 1, dtype=np.int64)
        return cls(df=df, nf1=nf1, nf2=nf2, sigma=5.0 * df, nwn=nwn)

    @classmethod
    def regional_fixed(cls, tlen: float, fmax: float, nwn: int, xlen: float, fmin: float = 0.0):
        """Regional grid: nwn wavenumbers k = (iwn-1) dwn, dwn = 2 pi / xlen, at every frequency
        (computeGreenFKSpectraForSynthetics.f90:195, 231-232)."""
        df = 1.0 / tlen
        nf2 = int(np.rint(fmax / df)) + 1
        nf1 = max(int(np.rint(fmin / df)), 2)
        g = cls(df=df, nf1=nf1, nf2=nf2, sigma=5.0 * df, nwn=np.full(nf2 - nf1 + 1, nwn, dtype=np.int64))
        g.dwn = MC_TWO_PID / xlen
        return g

    @classmethod
    def global_from_slowness(cls, tlen, fmax, p1lim, p2lim, frac, rearth, fmin=0.0):
        """nwn(jf) as in computeGreenFKSpectraForSynthetics.f90:221-236 (global mode)."""
        df = 1.0 / tlen
        nf2 = int(np.rint(fmax / df)) + 1
        nf1 = max(int(np.rint(fmin / df)), 2)
        ommax = MC_TWO_PID * (nf2 - 1) * df
        wnmax = 0.5 * ommax * (p1lim + p2lim) / (1.0 - frac)
        out = []
        for jf in range(nf1, nf2 + 1):
            f = (jf - 1) * df
            omre = 2.0 * MC_PI * f
            if f <= 0.5 * (nf2 - 1) * df:
                wn2 = omre * p1lim + frac * wnmax
            else:
                wn2 = 0.5 * ommax * (p1lim + p2lim) + (omre - ommax) * p2lim + frac * wnmax
            out.append(int(np.rint(wn2 * rearth - 0.5)) + 1)
        return cls(df=df, nf1=nf1, nf2=nf2, sigma=5.0 * df, nwn=np.asarray(out, dtype=np.int64))


def dll1_of(l: int, rearth: float) -> float:
    """dll1 = (wn*R)**2 with wn = sqrt(l*(l+1))/R (computeGreenFKSpectraForSynthetics.f90:413,
    geminiIntegrationEnvironment.f90:71)."""
    wn = math.sqrt(float(l * (l + 1))) / rearth
    return (wn * rearth) * (wn * rearth)
