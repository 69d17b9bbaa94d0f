"""Shared helpers for the tests: hand one scene to the engine and to the oracle."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc


def to_oracle(scene, threads: int = 0, integrator: int = orc.EULER) -> orc.Scene:
    """Concatenate a scene's bodies into one oracle scene (bodies are
    independent, so concatenation with index offsets changes nothing)."""
    pos, vel, m, T, ext, fixed, drag = [], [], [], [], [], [], []
    L, R, k, rest, damp, typ, om, aa, ph = [], [], [], [], [], [], [], [], []
    lc_ptr, lc_type, lc_vec, lc_num = [0], [], [], []
    any_ext = any(b.get("ext") is not None for b in scene.bodies)
    any_lc = any(b.get("lc_ptr") is not None for b in scene.bodies)
    off = 0
    for b in scene.bodies:
        n = len(b["pos"])
        s = len(b.get("left", ()))
        pos.append(np.asarray(b["pos"]).reshape(-1, 3))
        vel.append(np.zeros((n, 3)) if b.get("vel") is None else np.asarray(b["vel"]).reshape(-1, 3))
        m.append(np.broadcast_to(np.asarray(b["m"], dtype=float), (n,)))
        T.append(np.zeros(n) if b.get("T") is None else np.asarray(b["T"]))
        ext.append(np.zeros((n, 3)) if b.get("ext") is None else np.asarray(b["ext"]).reshape(-1, 3))
        fixed.append(np.zeros(n, np.uint8) if b.get("fixed") is None else np.asarray(b["fixed"], np.uint8))
        drag.append(np.zeros(n, np.int32) if b.get("drag") is None else np.asarray(b["drag"], np.int32))
        if s:
            l = np.asarray(b["left"], np.int64)
            r = np.asarray(b["right"], np.int64)
            L.append(np.where(l >= 0, l + off, -1))
            R.append(np.where(r >= 0, r + off, -1))

            def bc(x, default=0.0):
                return np.broadcast_to(np.asarray(default if x is None else x, dtype=float), (s,))

            k.append(bc(b["k"]))
            rest.append(bc(b["rest"]))
            damp.append(bc(b.get("damping")))
            typ.append(np.broadcast_to(np.asarray(0 if b.get("type") is None else b["type"]), (s,)))
            om.append(bc(b.get("omega")))
            aa.append(bc(b.get("act_a")))
            ph.append(bc(b.get("act_phi")))
        if any_lc:
            if b.get("lc_ptr") is None:
                lc_ptr.extend([lc_ptr[-1]] * n)
            else:
                p = np.asarray(b["lc_ptr"], np.int64)
                lc_ptr.extend(list(p[1:] + lc_ptr[-1]))
                lc_type.append(np.asarray(b["lc_type"], np.int32))
                lc_vec.append(np.asarray(b["lc_vec"], float).reshape(-1, 3))
                lc_num.append(np.asarray(b["lc_num"], float))
        off += n
    cat = np.concatenate
    kw = {}
    if L:
        kw = dict(left=cat(L), right=cat(R), k=cat(k), rest=cat(rest), damping=cat(damp),
                  type=cat(typ), omega=cat(om), act_a=cat(aa), act_phi=cat(ph))
    if any_lc:
        kw.update(lc_ptr=np.asarray(lc_ptr, np.int64),
                  lc_type=cat(lc_type) if lc_type else np.zeros(0, np.int32),
                  lc_vec=cat(lc_vec) if lc_vec else np.zeros((0, 3)),
                  lc_num=cat(lc_num) if lc_num else np.zeros(0))
    return orc.Scene(cat(pos), cat(m), vel=cat(vel), T=cat(T), ext=cat(ext) if any_ext else None,
                     fixed=cat(fixed), drag=cat(drag), planes=scene.planes, balls=scene.balls,
                     g=scene.g, dt=scene.dt, threads=threads, integrator=integrator, **kw)


def normwise_rel(a, b) -> float:
    """||a - b||_inf / ||b||_inf (velocities start at zero, so component-wise
    relative error is ill-posed)."""
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))
