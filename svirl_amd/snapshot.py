"""Checkpoint and resume of a GLSolver's state as a .npz file.

Saved: the geometry and dtype (checked on load), the order parameter and
vector potential (flat device layouts), the fixed-vortex potential vpi
when set (the TD steps change it, td.py:256-325), the TD solver's random
counter (TD._random_t, td.py:35-37) and the CG solver's hidden state that
the reference carries between cg() calls (search directions and betas,
cg.py:54-71). Loading into a solver built with the same arguments
continues a run bit for bit. Only numeric arrays are written, and the
loader never unpickles.
"""
import numpy as np

FORMAT = 1


def save(gl, path):
    """Write gl's state to `path` (.npz)."""
    c = gl.cfg
    psi = gl.vars._psi
    vp = gl.vars._vp
    psi.sync()
    vp.sync()
    d = {"format": np.int64(FORMAT), "Nx": np.int64(c.Nx), "Ny": np.int64(c.Ny),
         "dx": np.float64(c.dx), "dy": np.float64(c.dy), "dtype": np.array(np.dtype(c.dtype).str),
         "psi": np.asarray(psi._data).copy(), "ab": np.asarray(vp._data).copy()}
    fv = gl.params.fixed_vortices
    if fv._vpi is not None:
        fv._vpi.sync()
        d["vpi"] = np.asarray(fv._vpi._data).copy()
    td = gl.solve._td
    if td is not None:
        d["td_random_t"] = np.uint32(td._random_t)
        d["td_sweeps"] = np.asarray(td.sweeps, dtype=np.int32).reshape(-1, 2)
    cg = gl.solve._cg
    if cg is not None:
        d.update(cg._state())
    np.savez(path, **d)


def load(gl, path):
    """Restore the state saved by save() into gl (same geometry and dtype)."""
    with np.load(path, allow_pickle=False) as f:
        c = gl.cfg
        if int(f["format"]) != FORMAT:
            raise ValueError(f"snapshot format {int(f['format'])} is not {FORMAT}")
        if (int(f["Nx"]), int(f["Ny"])) != (int(c.Nx), int(c.Ny)) or \
                not np.isclose(float(f["dx"]), float(c.dx)) or not np.isclose(float(f["dy"]), float(c.dy)):
            raise ValueError("snapshot geometry does not match the solver")
        if str(f["dtype"]) != np.dtype(c.dtype).str:
            raise ValueError("snapshot dtype does not match the solver")
        fv = gl.params.fixed_vortices
        if ("vpi" in f.files) != (fv._vpi is not None):
            raise ValueError("snapshot and solver differ in having fixed vortices")
        psi = gl.vars._psi
        vp = gl.vars._vp
        psi.sync()
        vp.sync()
        psi._data[:] = f["psi"]
        psi.need_htod_sync()
        psi.sync()
        vp._data[:] = f["ab"]
        vp.need_htod_sync()
        vp.sync()
        if fv._vpi is not None:
            fv._vpi.sync()
            fv._vpi._data[:] = f["vpi"]
            fv._vpi.need_htod_sync()
            fv._vpi.sync()
        if "td_random_t" in f.files:
            gl.solve._init_td()
            gl.solve._td._random_t = np.uint32(f["td_random_t"])
            gl.solve._td.sweeps = [tuple(x) for x in f["td_sweeps"].tolist()]
        if "cg_dir_psi" in f.files:
            gl.solve._init_cg()
            gl.solve._cg._set_state({k: f[k] for k in f.files if k.startswith("cg_")})
    gl.solve.vortices_detected = False
