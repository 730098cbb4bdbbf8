"""Pins of the oracle's concentration solve (do_h_conc_solve,
genesis/src/hines/hines_conc.c:39-396) and concentration program
(do_h4_conc_chip_update, hines_conc_chip.c:59-222).

The solve is checked against a dense solve of the equations the reference
states in its header comment (hines_conc.c:58-103): shells coupled to their
neighbours by the covals0 values, buffer rows linearised around the
concentrations at t.  That matrix is assembled here row by row from the
model's arrays and solved with numpy, so an index, sign or ordering slip in
the elimination fails it.  Where the reference's elimination leaves two
fill-in terms out (diffusible buffers in shells past the first,
hines_conc.c:228-250 and :285-290), the residual of the dense system is
non-zero in exactly those rows and nowhere else.  Conservation of calcium
(free, and free plus bound) and fixed points (uniform concentration,
buffer equilibrium, a pump's Ceq) pin what the dense system alone does not:
the coupling values the setup writes."""
import numpy as np
import pytest

from hsolve_hip import model as M
from oracle import Oracle

RNG = np.random.default_rng(1234)


def couplings(m, i, l):
    """Split system i's covals0 values (cursor l) per shell and row, in the
    order the solve consumes them: fixed buffers' totals, then per
    diffusible buffer and the shell (down, up) with the missing ones None."""
    ms, dw, nd = int(m.nshells[i]), int(m.nshrows[i]), int(m.nshfix[i])
    cv = m.covals0
    out = []
    for s in range(ms):
        shell = {"btot": [], "dif": [], "shell": None}
        for n in range(nd):
            shell["btot"].append(cv[l]); l += 1
        def dnup():
            nonlocal l
            if s == 0:
                v = (None, cv[l]); l += 1
            elif s == ms - 1:
                v = (cv[l], None); l += 1
            else:
                v = (cv[l], cv[l + 1]); l += 2
            return v
        for n in range(nd, dw - 1):
            shell["dif"].append(dnup())
        shell["shell"] = dnup()
        out.append(shell)
    return out, l


def dense_system(m, i, l, conc, codiag, cores):
    """The dense matrix and right-hand side of system i (hines_conc.c:58-69),
    rows r = shell*dw + row; returns (A, b, l after the system)."""
    ms, dw, nd = int(m.nshells[i]), int(m.nshrows[i]), int(m.nshfix[i])
    c0, nb0 = int(m.dconc[i]), int(m.fbufval[i])
    cp, l = couplings(m, i, l)
    N = ms * dw
    A = np.zeros((N, N))
    b = np.zeros(N)
    C = conc[c0:c0 + N]
    bv = m.bufvals
    for s in range(ms):
        rS = s * dw + dw - 1
        ca = C[rS]
        # shell row
        dn, up = cp[s]["shell"]
        dn = dn or 0.0
        up = up or 0.0
        A[rS, rS] = codiag[c0 + rS] + dn + up
        if s > 0:
            A[rS, rS - dw] = -dn
        if s < ms - 1:
            A[rS, rS + dw] = -up
        if dw == 1:
            # shells alone: implicit in the half or whole step, right-hand
            # side C + coresult (hines_conc.c:322-393)
            b[rS] = ca + cores[c0 + rS]
            continue
        # with buffers: the Crank-Nicolson form, explicit half on the right
        b[rS] = cores[c0 + rS] + (1.0 - dn - up) * ca
        if s > 0:
            b[rS] += dn * C[rS - dw]
        if s < ms - 1:
            b[rS] += up * C[rS + dw]
        for n in range(dw - 1):
            r = s * dw + n
            f, kb, bt = bv[nb0 + 3 * n], bv[nb0 + 3 * n + 1], bv[nb0 + 3 * n + 2]
            B = C[r]
            tot = cp[s]["btot"][n] if n < nd else bt
            # buffer row: [1 + kb + f Ca] B* + f B Ca* = (1 - kb) B + tot (+ diffusion)
            A[r, r] = 1.0 + kb + f * ca
            A[r, rS] = f * B
            b[r] = B - kb * B + tot
            # the shell row's buffer terms
            A[rS, rS] += f * B
            A[rS, r] = kb + f * ca
            b[rS] += -kb * B + tot
            if n >= nd:
                bdn, bup = cp[s]["dif"][n - nd]
                bdn = bdn or 0.0
                bup = bup or 0.0
                A[r, r] += bdn + bup
                b[r] += -(bdn + bup) * B
                if s > 0:
                    A[r, r - dw] = -bdn
                    b[r] += bdn * C[r - dw]
                if s < ms - 1:
                    A[r, r + dw] = -bup
                    b[r] += bup * C[r + dw]
    return A, b, l


def solve_dense(m, conc, codiag, cores):
    """Every system's new concentrations from the dense equations."""
    out = conc.copy()
    l = 0
    for i in range(m.ndiffs):
        A, b, l = dense_system(m, i, l, conc, codiag, cores)
        x = np.linalg.solve(A, b)
        c0, N = int(m.dconc[i]), A.shape[0]
        if int(m.nshrows[i]) == 1 and m.method == M.CRANK:
            x = 2.0 * x - conc[c0:c0 + N]
        out[c0:c0 + N] = x
    return out


def shells_model(feats, method, shells=4, n=6, seed=0):
    return M.shell_population(n, (1, 3), (0.5, 0.5), shells=shells, features=feats, method=method,
                              seed=seed)


def perturbed(m, scale=0.3):
    """The model's concentrations with random relative perturbations."""
    return m.conc * (1.0 + scale * RNG.uniform(-1.0, 1.0, m.conc.size))


def system_rows(m, i):
    c0 = int(m.dconc[i])
    return c0, int(m.nshells[i]) * int(m.nshrows[i])


@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
@pytest.mark.parametrize("shells", [2, 3, 5])
def test_shells_alone_match_dense_solve(method, shells):
    m = shells_model(("ctaupump",), method, shells=shells)
    o = Oracle(m, givals=False)
    c = perturbed(m)
    o.put_conc(c)
    o.conc_solve()
    got = o.conc()
    want = solve_dense(m, c, np.ones(c.size), np.zeros(c.size))
    assert np.allclose(got, want, rtol=1e-12, atol=0), np.max(np.abs(got / want - 1))
    # the solve sets every diagonal it used back to 1
    d, _ = o.conc_regs()
    assert np.all(d == 1.0)


@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
def test_shells_alone_with_pumps_and_flux_match_dense_solve(method):
    """The solve takes the concops pass's diagonal and result."""
    m = shells_model(("ctaupump", "taupump", "mmpump", "hillpump", "leak", "kc"), method, shells=4)
    o = Oracle(m, givals=False)
    o.step(30)
    o.chip_update()
    o.conc_chip_update()
    c = o.conc()
    d, r = o.conc_regs()
    assert np.any(d != 1.0) and np.any(r != 0.0)
    o.conc_solve()
    want = solve_dense(m, c, d, r)
    assert np.allclose(o.conc(), want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
@pytest.mark.parametrize("shells", [3, 4, 6])
def test_shell_diffusion_conserves_calcium(method, shells):
    """sum(vol * C) is unchanged by pure diffusion (no flux, no pumps): the
    couplings of neighbouring shells agree (vol_m Dup_m = vol_m+1 Ddn_m+1).
    With two shells the reference's couplings differ (2 t0 + t1 against
    t0 + t1, hines_conc_init.c:362-387), checked below."""
    m = shells_model((), method, shells=shells)
    o = Oracle(m, givals=False)
    c = perturbed(m, 0.9)
    o.put_conc(c)
    o.conc_solve()
    got = o.conc()
    vol = m.extra["shell_vol"]
    for i in range(m.ndiffs):
        c0, N = system_rows(m, i)
        before = float(np.dot(vol[i], c[c0:c0 + N]))
        after = float(np.dot(vol[i], got[c0:c0 + N]))
        assert abs(after - before) <= 1e-14 * abs(before)


def test_two_shells_couple_asymmetrically():
    m = shells_model((), M.BEULER, shells=2)
    vol = m.extra["shell_vol"]
    for i in range(m.ndiffs):
        up0, dn1 = m.covals0[2 * i], m.covals0[2 * i + 1]
        assert vol[i][0] * up0 < vol[i][1] * dn1     # 2 t0 + t1 > t0 + t1


@pytest.mark.parametrize("method", [M.BEULER, M.CRANK])
def test_uniform_concentration_is_a_fixed_point(method):
    m = shells_model((), method, shells=5)
    o = Oracle(m, givals=False)
    c = np.full(m.conc.size, 7.5e-5)
    o.put_conc(c)
    o.conc_solve()
    assert np.allclose(o.conc(), c, rtol=1e-14, atol=0)


def test_fixed_buffers_match_dense_solve():
    m = shells_model(("fixbuf", "ctaupump", "mmpump", "leak", "kc"), M.CRANK, shells=4)
    o = Oracle(m, givals=False)
    o.step(20)
    o.chip_update()
    o.conc_chip_update()
    c = o.conc()
    d, r = o.conc_regs()
    o.conc_solve()
    want = solve_dense(m, c, d, r)
    assert np.allclose(o.conc(), want, rtol=1e-12, atol=0)


def test_fixed_buffers_conserve_total_calcium():
    """Free minus free buffer (= free + bound - total) summed over the
    shells' volumes is unchanged by diffusion and binding."""
    m = shells_model(("fixbuf",), M.CRANK, shells=4)
    o = Oracle(m, givals=False)
    c = perturbed(m, 0.5)
    o.put_conc(c)
    o.conc_solve()
    got = o.conc()
    vol = m.extra["shell_vol"]
    for i in range(m.ndiffs):
        c0, _ = system_rows(m, i)
        ca = lambda x: x[c0 + 1:c0 + 8:2]
        bf = lambda x: x[c0:c0 + 8:2]
        before = float(np.dot(vol[i], ca(c) - bf(c)))
        after = float(np.dot(vol[i], ca(got) - bf(got)))
        assert abs(after - before) <= 1e-13 * float(np.dot(vol[i], bf(c)))


def test_buffer_equilibrium_is_a_fixed_point():
    """Uniform Ca with the buffers at equilibrium stays put with fixed
    buffers; with a diffusible one the terms the reference's elimination
    leaves out move it slightly (the exact system would keep it)."""
    m = shells_model(("fixbuf",), M.CRANK, shells=4)
    o = Oracle(m, givals=False)
    c = m.conc.copy()      # uniform Ca, buffers at their equilibrium
    o.put_conc(c)
    o.conc_solve()
    assert np.allclose(o.conc(), c, rtol=1e-14, atol=0)
    m = shells_model(("fixbuf", "difbuf"), M.CRANK, shells=4)
    o = Oracle(m, givals=False)
    c = m.conc.copy()
    o.put_conc(c)
    o.conc_solve()
    dev = np.max(np.abs(o.conc() / c - 1.0))
    assert 1e-9 < dev < 1e-3
    x = solve_dense(m, c, np.ones(c.size), np.zeros(c.size))
    assert np.allclose(x, c, rtol=1e-12, atol=0)


def test_diffusible_buffers_deviate_only_where_the_reference_drops_terms():
    """With a diffusible buffer the reference's elimination leaves out the
    fill-in -H c2'' of the buffer row's shell entry (its first value is
    overwritten, hines_conc.c:228-250) and uses g5 for a5-corrected g5 in the
    shell row's next-shell entry (:285-290), both from the second shell on.
    Its result then solves the dense system exactly in every other row."""
    m = shells_model(("fixbuf", "difbuf", "ctaupump", "leak", "kc"), M.CRANK, shells=5)
    o = Oracle(m, givals=False)
    o.step(20)
    o.chip_update()
    o.conc_chip_update()
    c = o.conc()
    d, r = o.conc_regs()
    o.conc_solve()
    got = o.conc()
    l = 0
    ms, dw, nd = 5, 3, 1
    for i in range(m.ndiffs):
        A, b, l = dense_system(m, i, l, c, d, r)
        c0, N = system_rows(m, i)
        x = got[c0:c0 + N]
        res = np.abs(A @ x - b) / (np.abs(A) @ np.abs(x) + np.abs(b))
        dropped = set()
        for s in range(1, ms):
            for n in range(nd, dw - 1):
                dropped.add(s * dw + n)          # buffer rows of shells past the first
            if s < ms - 1:
                dropped.add(s * dw + dw - 1)     # shell rows between the first and last
        keep = [j for j in range(N) if j not in dropped]
        assert np.max(res[keep]) < 1e-14, (i, res)
        assert np.max(res[sorted(dropped)]) > 1e-12      # the terms are really left out
        # and the result stays near the exact solution
        want = np.linalg.solve(A, b)
        assert np.allclose(x, want, rtol=1e-3)


def test_ctaupump_relaxes_shells_to_ceq():
    """A constant-rate taupump alone (no flux) drives the pumped shell, and by
    diffusion the system, to its Ceq = 4e-5 mM."""
    m = shells_model(("ctaupump",), M.BEULER, shells=3)
    o = Oracle(m, givals=False)
    # the concentration passes alone: without the chip pass the shells'
    # flux stays at its initial 0
    o.put_conc(np.full(m.conc.size, 4e-5))
    o.conc_chip_update()
    o.conc_solve()
    assert np.allclose(o.conc(), 4e-5, rtol=1e-14)
    o.put_conc(np.full(m.conc.size, 2e-4))
    prev = np.inf
    for _ in range(40):
        for _ in range(100):
            o.conc_chip_update()
            o.conc_solve()
        dev = np.max(np.abs(o.conc() - 4e-5))
        assert dev < prev
        prev = dev
    assert np.allclose(o.conc(), 4e-5, rtol=1e-6)


def test_pool_and_pool_buffer_conserve_pool_calcium():
    """BUFFER_OP adds the buffer's change to the pool (hines_conc_chip.c:
    175-186); POOL_OP applies it: pool - free buffer is unchanged when nothing
    else feeds the pool."""
    m = shells_model(("pool",), M.BEULER, shells=3)
    o = Oracle(m, givals=False)
    # the pool sits after each compartment's system, its buffer after it
    kp = np.array([int(m.dconc[i]) + 3 for i in range(m.ndiffs)])
    c = m.conc.copy()
    c[kp + 1] *= 1.5        # free buffer off equilibrium
    o.put_conc(c)
    # shell Ca at 0: the Hill pump into the pool moves nothing
    cz = o.conc()
    for i in range(m.ndiffs):
        cz[int(m.dconc[i]):int(m.dconc[i]) + 3] = 0.0
    o.put_conc(cz)
    o.conc_chip_update()
    got = o.conc()
    assert np.allclose(got[kp] - got[kp + 1], cz[kp] - cz[kp + 1], rtol=1e-14)
    assert np.all(got[kp + 1] != cz[kp + 1])


def test_hill_pump_diagonal_closed_form():
    """HILLPUMP_OP adds dt vmax/vol * C^(h-1)/(C^h + Kd) to the shell's
    diagonal (hines_conc_chip.c:121-139), for Hill coefficients 1..4."""
    for hill in (1, 2, 3, 4):
        m = M.shell_population(3, (1,), (1.0,), shells=3, features=("hillpump",), method=M.BEULER,
                               hill=hill)
        o = Oracle(m, givals=False)
        o.conc_chip_update()
        d, _ = o.conc_regs()
        C = m.conc
        q = 0
        ops, cc = m.concops, m.concchip
        j = 0
        while ops[q] not in (10102, 10202):
            code = int(ops[q])
            if code in (10100, 10101):
                k = int(ops[q + 1]); q += 2
            elif code == 10103:
                q += 1; j += 2
            elif code == 11011:
                h = int(ops[q + 1]); q += 2
                Kd, a = cc[j], cc[j + 1]; j += 2
                x = C[k]
                want = 1.0 + a * x ** (h - 1) / (x ** h + Kd)
                assert h == hill
                assert d[k] == pytest.approx(want, rel=1e-14)
            else:
                raise AssertionError(code)


def test_beuler_buffers_are_refused_by_the_setup():
    with pytest.raises(ValueError):
        M.shell_population(2, (1,), (1.0,), features=("fixbuf",), method=M.BEULER)


def test_single_shell_systems_are_not_stepped():
    m = shells_model((), M.BEULER, shells=2)
    m.nshells[:] = 1
    o = Oracle(m, givals=False)
    with pytest.raises(RuntimeError):
        o.conc_solve()


def test_electrogenic_pump_closed_form():
    """ELMMPUMP_OP (hines_chip.c:278-302) after the CONC_VAL_OP of a shell:
    the current vmax*val*F * C/(C + Kd) joins the compartment's ichan (so
    results[2i] grows by it times tbyc) and dt*vmax/vol/(C + Kd) joins the
    shell's codiag, C the shell's concentration before the step."""
    base = M.shell_population(8, (1,), (1.0,), shells=3, features=("ctaupump",), method=M.BEULER, seed=4)
    m = M.shell_population(8, (1,), (1.0,), shells=3, features=("ctaupump", "elmmpump"), method=M.BEULER, seed=4)
    o0, o1 = Oracle(base, givals=False), Oracle(m, givals=False)
    o0.chip_update()
    o1.chip_update()
    r0, r1 = o0.state()["results"], o1.state()["results"]
    d1, _ = o1.conc_regs()
    ELM = 500
    for r in range(m.ncompts):
        a, b = int(m.compops[r]), int(m.compops[r + 1]) if r + 1 < m.ncompts else m.ops.size - 1
        q = a + int(np.flatnonzero(m.ops[a:b] == ELM)[0])
        # the chip values ELMMPUMP consumes: the pump's three, found after the
        # Ca channel's (Ek, gmax, m) in the compartment's chip run
        c0 = int(m.compchips[r])
        cend = int(m.compchips[r + 1]) if r + 1 < m.ncompts else m.chip.size
        Kd, I0, B0 = m.chip[cend - 5], m.chip[cend - 4], m.chip[cend - 3]
        k = int(m.ops[q - 2])                       # CONC_VAL_OP operand: the outermost shell
        C = m.conc[k]
        Ik = I0 * C * (1.0 / (C + Kd))
        tbyc = m.chip[cend - 2]
        assert r1[2 * r] - r0[2 * r] == pytest.approx(Ik * tbyc, rel=1e-9, abs=1e-18)
        assert d1[k] - 1.0 == pytest.approx(B0 / (C + Kd), rel=1e-12)
        assert r1[2 * r + 1] == r0[2 * r + 1]      # no conductance
