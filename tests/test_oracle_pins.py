"""CPU pins of the oracle functions that the other oracle tests leave open: the interpolated
bounce-back of f and g on non-uniform post-collision fields (Bouzidi, Firdaouss & Lallemand,
Phys. Fluids 13, 3452 (2001)), the new-fluid predicate against brute force, Nu with particle
nodes, and the output fields compute_Ked / compute_ted / overwrite_particle_nodes with the means
and profile built on them, all against closed forms.

Every test also checks that the plausible mistakes it is meant to catch (a swapped node or
population, a dropped factor, a strict inequality, the wrong geometry generation) would change
the result by far more than its tolerance, so a pass is not an accident of the chosen inputs.
"""
import numpy as np
import pytest

from cases import single_phase
from test_oracle import C, OPP, W

RT = 1.0 / 3.0


def sphere_case(L, R, npart, **kw):
    d = dict(single_phase(*L, beta=0.0), activate_particles=1, npart_override=npart,
             particle_diam_frac=2.0 * R / L[0], rho_particle_ratio=2.0)
    d.update(kw)
    return d


def node_xyz(L, idx):
    iz, rem = np.divmod(idx, L[0] * L[1])
    iy, ix = np.divmod(rem, L[0])
    return np.stack([ix, iy, iz], axis=-1)


def inside(L, x):
    return np.all((x >= 0) & (x < np.array(L)), axis=-1)


def at(L, x):
    return (x[..., 2] * L[1] + x[..., 1]) * L[0] + x[..., 0]


def linear_fields(rng, L, scale_a=1e-2, scale_b=1e-3):
    """F_i(x) = a_i + b_i . x with 27 distinct (a_i, b_i): f*_i differs from f*_opp and from
    f*_i at the neighbours."""
    n = L[0] * L[1] * L[2]
    xs = node_xyz(L, np.arange(n)).astype(np.float64)
    a = rng.normal(size=27) * scale_a
    b = rng.normal(size=(27, 3)) * scale_b
    return a, b, a[:, None] + b @ xs.T


def link_arrays(ora, L):
    lk = ora.links()
    x = np.stack([lk["cell_i"], lk["cell_j"], lk["cell_k"]], axis=1).astype(np.int64)
    return lk, x, lk["dir"].astype(np.int64), lk["q"].astype(np.float64), lk["pid"].astype(np.int64)


def wall_velocity(ora, x, i, q, pid):
    """u_w = V + Omega x (x_b - p) at the boundary point x_b = x + 1/2 + q c_i (ibb.cu:33-45)."""
    P = {k: ora.particles(k) for k in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z",
                                       "omega_x", "omega_y", "omega_z")}
    p = np.stack([P["pos_x"][pid], P["pos_y"][pid], P["pos_z"][pid]], axis=1)
    V = np.stack([P["vel_x"][pid], P["vel_y"][pid], P["vel_z"][pid]], axis=1)
    Om = np.stack([P["omega_x"][pid], P["omega_y"][pid], P["omega_z"][pid]], axis=1)
    xb = x + 0.5 + q[:, None] * C[i]
    return V + np.cross(Om, xb - p)


# Two spheres 1.4 nodes apart and one touching the x = 0 wall: the links cover all three
# branches of the bounce-back (upstream fluid with q below and above 1/2, upstream solid between
# the two spheres, upstream outside the domain at the wall).
L_IBB = (22, 16, 16)
R_IBB = 2.7
POS_IBB = np.array([[8.3, 7.6, 8.1], [14.1, 8.2, 7.7], [2.1, 8.4, 8.3]])


def ibb_setup(O, seed):
    ora = O.Oracle(O.default_case(**sphere_case(L_IBB, R_IBB, 3)))
    assert ora.derived.particle_radius == pytest.approx(R_IBB)
    ora.init(ppos=POS_IBB.T.reshape(-1), ptemp=np.array([0.31, -0.17, 0.05]))
    rng = np.random.default_rng(seed)
    for k, v in (("vel_x", [0.011, -0.006, 0.004]), ("vel_y", [-0.003, 0.008, 0.0]),
                 ("vel_z", [0.005, 0.002, -0.007]), ("omega_x", [0.002, -0.001, 0.003]),
                 ("omega_y", [-0.004, 0.002, 0.001]), ("omega_z", [0.001, 0.003, -0.002])):
        ora.set_particles(k, v)
    return ora, rng


def test_ibb_f_bouzidi_linear_fields(O):
    """bounce_back_particles_kernel (ibb.cu:5-82) on post-collision fields linear in position,
    F_i(x) = a_i + b_i.x.  Bouzidi's linear rule for a link (x, i, q) to a wall moving with u_w:
      q <  1/2 (x - c_i fluid): f_opp(x) = F_i(x - (1 - 2q) c_i) - mt, the post-collision value
                                interpolated at the point that reaches the wall in one step;
      q >= 1/2 (x - c_i fluid): f_opp(x) = [F_i(x) + (2q - 1) F_opp(x) - mt] / (2q);
      x - c_i solid or outside: f_opp(x) = F_i(x) - mt;
    with mt = 2 w_i RHO0 (c_i . u_w) / RT.  For a linear field the first rule is exact."""
    ora, rng = ibb_setup(O, 21)
    a, b, F = linear_fields(rng, L_IBB)
    ora.set_populations("f_collide", F)
    ora.stage("ibb_f")
    out = ora.populations("f")
    ib, _ = ora.flags()
    lk, x, i, q, pid = link_arrays(ora, L_IBB)
    opp = OPP[i]
    uw = wall_velocity(ora, x, i, q, pid)
    mt = 2.0 * W[i] * 1.0 * np.einsum("ka,ka->k", C[i], uw) / RT
    up = x - C[i]
    up_fluid = inside(L_IBB, up)
    up_fluid[up_fluid] = ib[at(L_IBB, up[up_fluid])] == 0
    lin = lambda pop, pts: a[pop] + np.einsum("ka,ka->k", b[pop], pts)
    Fi_x, Fopp_x = lin(i, x), lin(opp, x)
    lo = up_fluid & (q < 0.5)
    hi = up_fluid & (q >= 0.5)
    el = ~up_fluid
    want = np.where(lo, lin(i, x - (1.0 - 2.0 * q)[:, None] * C[i]) - mt,
                    np.where(hi, (Fi_x + (2.0 * q - 1.0) * Fopp_x - mt) / (2.0 * q), Fi_x - mt))
    got = out[opp, at(L_IBB, x)]
    assert lo.sum() > 50 and hi.sum() > 50 and el.sum() > 5, (lo.sum(), hi.sum(), el.sum())
    assert (el & inside(L_IBB, up)).sum() > 0 and (el & ~inside(L_IBB, up)).sum() > 0
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)
    # the mistakes this catches, each far outside the tolerance
    swapped_nodes = lin(i, x - (2.0 * q)[:, None] * C[i]) - mt           # x <-> x - c_i
    swapped_pops = lin(opp, x - (1.0 - 2.0 * q)[:, None] * C[i]) - mt    # i <-> opp (q < 1/2)
    swapped_hi = (Fopp_x + (2.0 * q - 1.0) * Fi_x - mt) / (2.0 * q)       # i <-> opp (q >= 1/2)
    no_mt = Fi_x
    for name, variant, sel in (("x<->x-c", swapped_nodes, lo), ("i<->opp", swapped_pops, lo),
                               ("i<->opp, q>=1/2", swapped_hi, hi), ("no mt", no_mt, el)):
        assert np.max(np.abs(variant[sel] - got[sel])) > 1e-5, name


def test_ibb_g_bouzidi_linear_fields_with_density(O):
    """bounce_back_thermal_kernel (ibb.cu:84-136) on linear post-collision fields G_i, a
    non-uniform density deviation rho(x) and distinct sphere temperatures T_p, with
    amp = 2 w_i T_p (1 + rho(x)) (the Dirichlet value of the anti-bounce-back):
      q <= 1/2 (x - c_i fluid in the previous geometry): g_opp = -G_i(x - (1 - 2q) c_i) + amp;
      q >  1/2:                                           g_opp = [-G_i(x) + amp + (2q-1) G_opp(x)] / (2q);
      otherwise:                                          g_opp = -G_i(x) + amp.
    The previous geometry is made to differ from the current one (the spheres are moved and
    build_links run), so the test also fixes which flags decide 'upstream fluid' (ibnode_prev,
    ibb.cu:116) and which the fallback (ibnode, ibb.cu:126)."""
    ora, rng = ibb_setup(O, 22)
    moved = POS_IBB + np.array([[0.6, -0.4, 0.3], [-0.5, 0.45, 0.2], [0.7, 0.1, -0.35]])
    for a_, k in enumerate(("pos_x", "pos_y", "pos_z")):
        ora.set_particles(k, moved[:, a_])
    ora.stage("build_links")
    ib, _ = ora.flags()
    ib_prev, _ = ora.flags(prev=True)
    assert np.any(ib != ib_prev)
    a, b, G = linear_fields(rng, L_IBB, scale_a=0.2, scale_b=0.01)
    rho = rng.normal(size=ib.size) * 0.05
    ora.set_populations("g_collide", G)
    ora.set_field("rho", rho)
    ora.stage("ibb_g")
    out = ora.populations("g")
    lk, x, i, q, pid = link_arrays(ora, L_IBB)
    opp = OPP[i]
    Tp = ora.particles("temp")[pid]
    idx = at(L_IBB, x)
    amp = 2.0 * W[i] * Tp * (1.0 + rho[idx])
    up = x - C[i]
    up_in = inside(L_IBB, up)
    up_fluid = up_in.copy()
    up_fluid[up_in] = ib_prev[at(L_IBB, up[up_in])] == 0
    lin = lambda pop, pts: a[pop] + np.einsum("ka,ka->k", b[pop], pts)
    Gi_x, Gopp_x = lin(i, x), lin(opp, x)
    lo = up_fluid & (q <= 0.5)
    hi = up_fluid & (q > 0.5)
    el = ~up_fluid
    want = np.where(lo, -lin(i, x - (1.0 - 2.0 * q)[:, None] * C[i]) + amp,
                    np.where(hi, (-Gi_x + amp + (2.0 * q - 1.0) * Gopp_x) / (2.0 * q),
                             -Gi_x + amp))
    got = out[opp, idx]
    assert lo.sum() > 50 and hi.sum() > 50 and el.sum() > 5, (lo.sum(), hi.sum(), el.sum())
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
    # upstream fluid by the current flags where the previous ones disagree
    cur_fluid = up_in.copy()
    cur_fluid[up_in] = ib[at(L_IBB, up[up_in])] == 0
    assert np.any(cur_fluid != up_fluid)
    amp_no_rho = 2.0 * W[i] * Tp
    amp_rho_up = 2.0 * W[i] * Tp * (1.0 + rho[at(L_IBB, np.where(up_in[:, None], up, x))])
    for name, variant, sel in (
            ("no (1+rho)", want - amp + amp_no_rho, np.ones_like(lo)),
            ("rho at x-c", want - amp + amp_rho_up, up_in),
            ("x<->x-c", -lin(i, x - (2.0 * q)[:, None] * C[i]) + amp, lo),
            ("i<->opp", -lin(opp, x - (1.0 - 2.0 * q)[:, None] * C[i]) + amp, lo),
            ("i<->opp, q>1/2", (-Gopp_x + amp + (2.0 * q - 1.0) * Gi_x) / (2.0 * q), hi)):
        assert np.max(np.abs(variant[sel] - got[sel])) > 1e-5, name


# New fluid: every sphere has radius R_NF = 1.75 + 2^-52 in a lattice 32 wide, so that
# 0.5 * (2 R / 32) * 32 returns R exactly and R + CLOSE_THRESH is exactly a double: a node at
# exactly that distance from a sphere's new centre sits on the threshold of refill.cu:39.
CLOSE_THRESH = 1.7606816861659007
R_NF = 1.75 + 2.0 ** -52
L_NF = (32, 18, 18)


def test_new_fluid_predicate_brute_force(O):
    """new_fluid_count_kernel / new_fluid_fill_kernel (refill.cu:8-76): a node is new fluid iff
    it was solid (ibnode_prev = 1), is fluid now (ibnode = 0), its previous owner is a valid pid,
    and its distance to that owner's NEW centre minus R is <= close_thresh; listed in idx order
    with the previous owner.  Brute force over every node of a 32x18x18 lattice where sphere 0
    jumps 3 nodes (so the old and new centres give different sets), sphere 1 moves 2 nodes and
    leaves a node exactly on the threshold, and sphere 2 does not move."""
    d = R_NF + CLOSE_THRESH
    assert d - R_NF == CLOSE_THRESH and np.sqrt(d * d) == d
    c = np.array([6.5, 8.5, 8.5])                     # the threshold node's centre
    p1_new = np.array([6.5 - d, 8.5, 8.5])
    assert 6.5 - p1_new[0] == d
    old = np.array([[18.0, 8.3, 8.6], p1_new + [2.0, 0.0, 0.0], [26.2, 12.4, 11.1]])
    new = np.array([[21.0, 8.3, 8.6], p1_new, [26.2, 12.4, 11.1]])
    kw = sphere_case(L_NF, R_NF, 3)
    ora = O.Oracle(O.default_case(**kw))
    assert ora.derived.particle_radius == R_NF
    ora.init(ppos=old.T.reshape(-1))
    for k, a_ in (("pos_x", 0), ("pos_y", 1), ("pos_z", 2)):
        ora.set_particles(k, new[:, a_])
    ora.stage("build_links")
    ib, ow = ora.flags()
    ib_prev, ow_prev = ora.flags(prev=True)
    ora.stage("refill")
    nodes, pids = ora.new_fluid()

    n = ib.size
    xc = node_xyz(L_NF, np.arange(n)).astype(np.float64) + 0.5

    def predicate(flags_prev, flags_now, owners, centres, strict=False):
        sel = (flags_prev == 1) & (flags_now == 0) & (owners >= 0) & (owners < 3)
        o = np.where(sel, owners, 0)
        dx = xc[:, 0] - centres[o, 0]
        dy = xc[:, 1] - centres[o, 1]
        dz = xc[:, 2] - centres[o, 2]
        dist = np.sqrt(dx * dx + dy * dy + dz * dz)
        close = (dist - R_NF < CLOSE_THRESH) if strict else (dist - R_NF <= CLOSE_THRESH)
        keep = sel & close
        return np.flatnonzero(keep), owners[keep]

    want_nodes, want_pids = predicate(ib_prev, ib, ow_prev, new)
    assert np.array_equal(nodes, want_nodes) and np.array_equal(pids, want_pids)
    assert set(np.unique(pids)) == {0, 1}
    k = int(at(L_NF, (c - 0.5).astype(np.int64)))
    assert k in set(nodes.tolist()) and pids[list(nodes).index(k)] == 1
    # the mistakes this catches
    variants = {
        "strict <": predicate(ib_prev, ib, ow_prev, new, strict=True)[0],
        "old centres": predicate(ib_prev, ib, ow_prev, old)[0],
        "current owner": predicate(ib_prev, ib, ow, new)[0],
        "flags swapped": predicate(ib, ib_prev, ow_prev, new)[0],
    }
    for name, v in variants.items():
        assert not np.array_equal(v, nodes), name


def test_nu_with_particle_nodes_closed_form(O):
    """output_nu (output_statis.cu:111-223) with particle nodes.  The fields are constant on the
    fluid and on the particle nodes of every x-plane, u_x phi = A_f(ix) and A_p(ix), so the plane
    averages are those constants and
      Nu_bulk_fluid = (1 - phi_p) (1 + <A_f>/q),   Nu_bulk_part = phi_p Cp <A_p>/q,
      q = kappa dT / LX,  <A> = (1/LX) sum over planes (planes without particle nodes add 0),
      wall Nu = (fluid nodes of the wall plane) * 2 |T_wall - phi_wall| / (LY LZ) * LX / dT,
    with phi_p the PARTICLE_FRACTION parameter, not the actual fraction, and the current
    geometry (the previous one is made to differ)."""
    L = (12, 10, 9)
    pf, cp = 0.0371, 2.5
    kw = sphere_case(L, 2.2, 2, particle_fraction=pf, cp_particle_ratio=cp, t_hot=0.5,
                     t_cold=-0.5, t_ref=0.0)
    ora = O.Oracle(O.default_case(**kw))
    ora.init(ppos=np.array([[1.1, 4.9, 4.6], [6.8, 5.2, 4.3]]).T.reshape(-1))
    for k, v in (("pos_x", [1.3, 7.4]), ("pos_y", [5.0, 4.6]), ("pos_z", [4.5, 4.1])):
        ora.set_particles(k, v)
    ora.stage("build_links")
    ib, _ = ora.flags()
    ib_prev, _ = ora.flags(prev=True)
    assert np.any(ib != ib_prev)
    n = ib.size
    ix = node_xyz(L, np.arange(n))[:, 0]
    phi_f = 0.5 - (ix + 0.5) / L[0] + 0.013 * np.sin(ix)
    u_f = 0.01 * np.cos(ix) + 0.004
    phi_p, u_p = 0.3 + 0.02 * ix, -0.02 + 0.001 * ix
    solid = ib == 1
    ora.set_field("phi", np.where(solid, phi_p, phi_f))
    ora.set_field("ux", np.where(solid, u_p, u_f))
    nu = ora.nu()
    LX, LY, LZ = L
    kappa = ora.derived.diff
    qflux = kappa * 1.0 / LX
    planes = np.arange(LX)
    Af = (0.01 * np.cos(planes) + 0.004) * (0.5 - (planes + 0.5) / LX + 0.013 * np.sin(planes))
    has_p = np.array([solid[ix == k].any() for k in planes])
    Ap = np.where(has_p, (-0.02 + 0.001 * planes) * (0.3 + 0.02 * planes), 0.0)
    assert has_p.sum() < LX and has_p[0]
    nb_fluid = (~solid[ix == 0]).sum()
    nt_fluid = (~solid[ix == LX - 1]).sum()
    assert 0 < nb_fluid < LY * LZ
    Tb = nb_fluid * 2.0 * (0.5 - phi_f[ix == 0][0]) / (LY * LZ) * LX / 1.0
    Tt = nt_fluid * 2.0 * (phi_f[ix == LX - 1][0] + 0.5) / (LY * LZ) * LX / 1.0
    nbf = (1 - pf) * (1 + Af.mean() / qflux)
    nbp = pf * cp * Ap.sum() / LX / qflux
    want = np.array([0.0, pf, nbf + nbp, Tt, Tb, nbf, nbp])
    np.testing.assert_allclose(nu, want, rtol=1e-12, atol=1e-14)
    # the mistakes this catches
    actual_frac = solid.mean()
    assert abs(pf * cp * Ap[has_p].mean() / qflux - nu[6]) > 1e-6 * abs(nu[6])
    assert abs((1 - actual_frac) * (1 + Af.mean() / qflux) - nu[5]) > 1e-6 * abs(nu[5])
    assert abs(Tb * LY * LZ / nb_fluid - nu[4]) > 1e-6 * abs(nu[4])
    prev_fluid_wall = (ib_prev[ix == 0] == 0).sum()
    assert prev_fluid_wall != nb_fluid


def M4(a, b, c, d):
    """sum_i w_i c_ia c_ib c_ic c_id of D3Q27 (fourth-order isotropy)."""
    return ((a == b) * (c == d) + (a == c) * (b == d) + (a == d) * (b == c)) / 9.0


def test_ked_ted_closed_form(O):
    """compute_Ked (collision.cu:91-152) and compute_ted (collision.cu:154-198) on populations
    built as equilibrium plus a known non-equilibrium part, with a force (F_y, F_z too):
      f* = feq(rho, u) + w_i c_i.A.c_i  ->  sum c_a c_b (f* - feq) = (d_ab tr A + 2 A_ab)/9
      sum c_a c_b (c.F - F.u) feq / (2 RT) = 1.5 [ (d_ab F.u + F_a u_b + F_b u_a)/3
                                     - F.u ((rho - 1.5 u^2) d_ab/3 + (d_ab u^2 + 2 u_a u_b)/2) ]
      tau_ab = -(1 - 1/(2 tau)) (both),  S = tau/(2 nu),  Ked = nu/2 (4 S_aa^2 + 8 S_ab^2);
      g* = geqnew(phi, rho, u) + w_i c_i.B  ->  gneq = B/3,
      t = -(2 gneq + phi F)/(2 tauc RT + rho RT),  ted = kappa |t|^2,
    from the D3Q27 moments (fourth-order isotropy, odd moments zero).  feq and geqnew are
    written from their definitions (collision.cu:16-23, 53-56)."""
    L = (7, 6, 5)
    ora = O.Oracle(O.default_case(**single_phase(*L, ra=1e6)))
    ora.init()
    n = ora.n
    rng = np.random.default_rng(31)
    rho = rng.normal(size=n) * 0.02
    u = rng.normal(size=(3, n)) * 0.02
    F = rng.normal(size=(3, n)) * 1e-3
    phi = 0.1 + rng.normal(size=n) * 0.05
    A = rng.normal(size=(3, 3, n)) * 1e-3
    A = 0.5 * (A + A.transpose(1, 0, 2))
    B = rng.normal(size=(3, n)) * 1e-3
    cu = C @ u
    u2 = (u * u).sum(0)
    feq = W[:, None] * (rho + 3.0 * cu + 4.5 * cu ** 2 - 1.5 * u2)
    geq = W[:, None] * phi * (1.0 + 3.0 * cu + 4.5 * cu ** 2 - 1.5 * u2)
    geq[0] -= (1.0 - W[0]) * phi * rho
    geq[1:] += W[1:, None] * phi * rho
    fneq = W[:, None] * np.einsum("ia,abn,ib->in", C, A, C)
    gneq_pop = W[:, None] * (C @ B)
    ora.set_populations("f_collide", feq + fneq)
    ora.set_populations("g_collide", geq + gneq_pop)
    for k, v in (("rho", rho), ("ux", u[0]), ("uy", u[1]), ("uz", u[2]), ("phi", phi),
                 ("fx", F[0]), ("fy", F[1]), ("fz", F[2])):
        ora.set_field(k, v)
    out = ora.output_fields()
    tau, tauc = ora.derived.tau, ora.derived.tauc
    visc, kappa = ora.case.visc, ora.derived.diff
    I3 = np.eye(3)
    trA = np.einsum("aan->n", A)
    Fu = (F * u).sum(0)
    stress = np.zeros((3, 3, n))
    for a in range(3):
        for b in range(3):
            neq = (I3[a, b] * trA + 2.0 * A[a, b]) / 9.0
            src = 1.5 * ((I3[a, b] * Fu + F[a] * u[b] + F[b] * u[a]) / 3.0
                         - Fu * ((rho - 1.5 * u2) * I3[a, b] / 3.0 + (I3[a, b] * u2 + 2.0 * u[a] * u[b]) / 2.0))
            stress[a, b] = -(1.0 - 0.5 / tau) * (neq + src)
    S = stress / (2.0 * visc)
    ked = 0.5 * visc * (4.0 * (S[0, 0] ** 2 + S[1, 1] ** 2 + S[2, 2] ** 2)
                        + 8.0 * (S[0, 1] ** 2 + S[0, 2] ** 2 + S[1, 2] ** 2))
    t = -(2.0 * B / 3.0 + phi * F) / (2.0 * tauc * RT + rho * RT)
    ted = kappa * (t * t).sum(0)
    np.testing.assert_allclose(out[6], ked, rtol=1e-9, atol=1e-22)
    np.testing.assert_allclose(out[5], ted, rtol=1e-11, atol=1e-22)
    for k, v in enumerate((rho, u[0], u[1], u[2], phi)):
        assert np.array_equal(out[k], v)
    # the mistakes this catches: the force term left out, the ted denominator with rho + 1
    no_force = 0.5 * visc * sum(((-(1.0 - 0.5 / tau) * (I3[a, b] * trA + 2.0 * A[a, b]) / 9.0)
                                 / (2.0 * visc)) ** 2 * (4.0 if a == b else 8.0)
                                for a in range(3) for b in range(a, 3))
    assert np.max(np.abs(no_force - out[6]) / np.abs(out[6])) > 1e-6
    t1 = -(2.0 * B / 3.0 + phi * F) / (2.0 * tauc * RT + (1.0 + rho) * RT)
    assert np.max(np.abs(kappa * (t1 * t1).sum(0) - out[5]) / out[5]) > 1e-3


def test_output_overwrite_means_profile_closed_form(O):
    """overwrite_particle_nodes (output_flow.cu:5-60 / output_statis.cu:5-60) replaces, at every
    particle node, rho by RHO0, u by the rigid-body velocity V + Omega x (x - p), phi by T_p and
    ted, Ked by 0; diag_flow (output_statis.cu:64-109) and output_profile
    (output_statis.cu:225-285) then average the overwritten fields over the lattice and over
    each x-plane."""
    L = (14, 12, 11)
    kw = sphere_case(L, 2.6, 2)
    ora = O.Oracle(O.default_case(**kw))
    P = np.array([[4.3, 6.1, 5.4], [9.6, 5.7, 5.9]])
    ora.init(ppos=P.T.reshape(-1), ptemp=np.array([0.4, -0.25]))
    V = np.array([[0.01, -0.02, 0.005], [-0.004, 0.003, 0.012]])
    Om = np.array([[0.002, 0.001, -0.003], [-0.001, 0.004, 0.002]])
    for a_, k in enumerate(("vel_x", "vel_y", "vel_z")):
        ora.set_particles(k, V[:, a_])
    for a_, k in enumerate(("omega_x", "omega_y", "omega_z")):
        ora.set_particles(k, Om[:, a_])
    n = ora.n
    rng = np.random.default_rng(41)
    fields = {k: rng.normal(size=n) * 0.01 for k in ("rho", "ux", "uy", "uz", "phi")}
    for k, v in fields.items():
        ora.set_field(k, v)
    ib, ow = ora.flags()
    solid = ib == 1
    assert solid.sum() > 50 and set(np.unique(ow[solid])) == {0, 1}
    x = node_xyz(L, np.arange(n)).astype(np.float64) + 0.5
    o = np.where(solid, ow, 0)
    urb = V[o] + np.cross(Om[o], x - P[o])
    want = np.stack([np.where(solid, 1.0, fields["rho"]),
                     np.where(solid, urb[:, 0], fields["ux"]),
                     np.where(solid, urb[:, 1], fields["uy"]),
                     np.where(solid, urb[:, 2], fields["uz"]),
                     np.where(solid, np.array([0.4, -0.25])[o], fields["phi"])])
    out = ora.output_fields()
    np.testing.assert_allclose(out[:5], want, rtol=0, atol=1e-16)
    assert np.all(out[5:, solid] == 0.0)
    np.testing.assert_allclose(ora.means(), want[1:].mean(axis=1), rtol=1e-13, atol=1e-17)
    ix = node_xyz(L, np.arange(n))[:, 0]
    prof = np.stack([[want[k][ix == p].mean() for k in (1, 2, 3, 4)] for p in range(L[0])])
    np.testing.assert_allclose(ora.profile(), prof, rtol=1e-13, atol=1e-17)
    # the mistake this catches: the particle's centre velocity without the rotation
    assert np.max(np.abs(V[o][solid, 0] - out[1, solid])) > 1e-4
