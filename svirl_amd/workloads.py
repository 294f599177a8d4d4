"""Seeded synthetic inputs for the benchmark configurations.

Config 1: 129^2 nodes, kappa = inf. Config 2: 1025^2 nodes, kappa = 5.
Config 3: a 4097^2 node instance (Lx = Ly = 2048, dx = 0.5) with a material
mask of random circular holes (radius U[4, 10], added until 5% of the cell
area is removed) and a Gaussian random pinning field eps(r) (mean 1, std
0.2, correlation length 2 = 4 grid spacings). Smaller Lx gives the same
statistics on a smaller grid. Configs 4 and 5: batches of independent
instances (config4_specs, config5_specs).
"""
import numpy as np


def holes_mask(Nxc, Nyc, dx, dy, seed=30, r_min=4.0, r_max=10.0, fraction=0.05):
    """Cell tiling (Nxc, Nyc) bool, False inside seeded random circles."""
    rng = np.random.default_rng(seed)
    mt = np.ones((Nxc, Nyc), dtype=bool)
    Lx, Ly = Nxc * dx, Nyc * dy
    target = int(round(fraction * Nxc * Nyc))
    removed = 0
    while removed < target:
        x0, y0 = rng.random() * Lx, rng.random() * Ly
        r = r_min + (r_max - r_min) * rng.random()
        i0, i1 = max(0, int((x0 - r) / dx)), min(Nxc, int((x0 + r) / dx) + 2)
        j0, j1 = max(0, int((y0 - r) / dy)), min(Nyc, int((y0 + r) / dy) + 2)
        if i0 >= i1 or j0 >= j1:
            continue
        xc = (np.arange(i0, i1) + 0.5) * dx
        yc = (np.arange(j0, j1) + 0.5) * dy
        inside = (xc[:, None] - x0) ** 2 + (yc[None, :] - y0) ** 2 < r * r
        sub = mt[i0:i1, j0:j1]
        removed += int(np.count_nonzero(sub & inside))
        sub[inside] = False
    return mt


def pinning_field(Nx, Ny, dx, seed=31, mean=1.0, std=0.2, corr_length=2.0):
    """Node field (Nx, Ny): Gaussian-filtered white noise, rescaled."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((Nx, Ny))
    f = gaussian_filter(w, sigma=corr_length / dx, mode='wrap')
    f = (f - f.mean()) / f.std()
    return mean + std * f


def config1(dtype=np.float64, L=64.0, dx=0.5, seed_psi=1, H=0.1):
    """Config 1 as a GLSolver: 129^2 nodes, kappa = inf (A fixed at the
    symmetric gauge of H), seeded random psi, no mask, eps = 1."""
    from svirl_amd import GLSolver
    return GLSolver(Lx=L, Ly=L, dx=dx, dy=dx, order_parameter='random', random_seed=seed_psi, random_level=1.0,
                    gl_parameter=np.inf, homogeneous_external_field=H, dtype=dtype)


def config2(dtype=np.float64, L=512.0, dx=0.5, seed_psi=2, kappa=5.0, sigma=200.0, H=0.1):
    """Config 2 as a GLSolver: 1025^2 nodes, kappa = 5, sigma = 200,
    H = 0.1, seeded random psi, no mask, eps = 1."""
    from svirl_amd import GLSolver
    return GLSolver(Lx=L, Ly=L, dx=dx, dy=dx, order_parameter='random', random_seed=seed_psi, random_level=1.0,
                    gl_parameter=kappa, normal_conductivity=sigma, homogeneous_external_field=H, dtype=dtype)


def config3(L=2048.0, dx=0.5, dtype=np.float64, seed_psi=3, kappa=5.0, sigma=200.0, H=0.05):
    """Config 3 as a GLSolver (mask, eps map, seeded random psi)."""
    from svirl_amd import GLSolver
    gl = GLSolver(Lx=L, Ly=L, dx=dx, dy=dx, order_parameter='random', random_seed=seed_psi,
                  gl_parameter=kappa, normal_conductivity=sigma, homogeneous_external_field=H, dtype=dtype)
    c = gl.cfg
    gl.mesh.material_tiling = holes_mask(int(c.Nxc), int(c.Nyc), float(c.dx), float(c.dy))
    gl.vars.set_order_parameter_to_zero_outside_material()
    gl.params.linear_coefficient = pinning_field(int(c.Nx), int(c.Ny), float(c.dx))
    return gl


def config4_specs(n=256, seed=40, sizes=(32, 64, 128, 256, 512), dx=0.5, sigma=200.0):
    """The per-instance draws of Config 4 (host only): Lx and Ly drawn
    independently from `sizes` (powers of two by default), H uniform in
    [0.02, 0.5], kappa from {2, 5, inf}, and the seed of the instance's
    random psi and pinning field eps(r) (mean 1, std 0.2, correlation 2)."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n):
        Lx, Ly = float(rng.choice(sizes)), float(rng.choice(sizes))
        H = float(rng.uniform(0.02, 0.5))
        kappa = float(rng.choice([2.0, 5.0, np.inf]))
        specs.append(dict(Lx=Lx, Ly=Ly, dx=dx, H=H, kappa=kappa, sigma=sigma, seed=seed * 1000 + k,
                          Nx=int(round(Lx / dx)) + 1, Ny=int(round(Ly / dx)) + 1))
    return specs


def config5_specs(n=1024, seed=50, L=128.0, dx=0.5, kappa=5.0):
    """The per-instance draws of Config 5 (host only): (2L/dx+1)^2 nodes
    (257^2 by default), kappa = 5, H uniform in [0.05, 0.3], the seed of the
    instance's random psi and pinning field eps(r) (std 0.2). sigma is the
    reference's default 1 (CG does not read it)."""
    rng = np.random.default_rng(seed)
    N = int(round(L / dx)) + 1
    return [dict(Lx=L, Ly=L, dx=dx, H=float(rng.uniform(0.05, 0.3)), kappa=kappa, sigma=1.0,
                 seed=seed * 10000 + k, Nx=N, Ny=N) for k in range(n)]


def instance(spec, dtype):
    """A GLSolver of one Config 4 / Config 5 draw."""
    from svirl_amd import GLSolver
    gl = GLSolver(Lx=spec["Lx"], Ly=spec["Ly"], dx=spec["dx"], dy=spec["dx"], order_parameter='random',
                  random_seed=spec["seed"], gl_parameter=spec["kappa"], normal_conductivity=spec["sigma"],
                  homogeneous_external_field=spec["H"], dtype=dtype)
    c = gl.cfg
    gl.params.linear_coefficient = pinning_field(int(c.Nx), int(c.Ny), spec["dx"], seed=spec["seed"])
    return gl


def config4(n=256, dtype=np.float32, seed=40, sizes=(32, 64, 128, 256, 512), dx=0.5, sigma=200.0):
    """Config 4: n independent instances for a batched run (config4_specs)."""
    return [instance(s, dtype) for s in config4_specs(n, seed, sizes, dx, sigma)]


def config5(n=1024, dtype=np.float32, seed=50, L=128.0, dx=0.5, kappa=5.0):
    """Config 5: n instances for a batched CG minimisation (config5_specs)."""
    return [instance(s, dtype) for s in config5_specs(n, seed, L, dx, kappa)]
