"""Mutation check of the oracle's pinning tests.

Each case plants one plausible mistake in a copy of oracle/femx_oracle.c (a
flipped sign, a dropped coefficient or term, a transposed index), builds the
copy with the oracle's own recipe (make -C oracle variant), and runs the
closed-form and invariant tests against it through FEMX_ORACLE_LIB: every
mutant must make at least one of them fail.  The planted mistakes live only
in a temporary directory, on the CPU."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CHECKS = [os.path.join(ROOT, "tests", f) for f in ("test_oracle_terms.py", "test_oracle_physics.py")]

MUTANTS = [
    ("vms acceleration sign", "lRes[a][i] += wGp * ((u_gp_curr[i] - u_gp_prev[i]) / dt) * VLN[g][a];", "lRes[a][i] -= wGp * ((u_gp_curr[i] - u_gp_prev[i]) / dt) * VLN[g][a];"),
    ("vms acceleration prev dropped", "((u_gp_curr[i] - u_gp_prev[i]) / dt)", "((u_gp_curr[i]) / dt)"),
    ("vms diffusion nu dropped", "lRes[a][i] += wGp * nu * (gradHdu[i][j]) * DN[j][a];", "lRes[a][i] += wGp * (gradHdu[i][j]) * DN[j][a];"),
    ("vms diffusion gradient transposed", "lRes[a][i] += wGp * nu * (gradHdu[i][j]) * DN[j][a];", "lRes[a][i] += wGp * nu * (gradHdu[j][i]) * DN[j][a];"),
    ("vms kinematic viscosity mu*rho", "  const double nu = s->p.mu / s->p.rho;\n  const double invEpsilon", "  const double nu = s->p.mu * s->p.rho;\n  const double invEpsilon"),
    ("vms convection sign", "lRes[a][i] += wGp * ahGradHu[i] * VLN[g][a];", "lRes[a][i] -= wGp * ahGradHu[i] * VLN[g][a];"),
    ("vms convection gradient transposed", "ahGradHu[i] += gradHdu[i][j] * hah[g][j];", "ahGradHu[i] += gradHdu[j][i] * hah[g][j];"),
    ("vms advection velocity without subgrid", "      hah[g][i] += sdu[g][i];\n", "\n"),
    ("vms pressure sign", "lRes[a][i] += -wGp * hph * DN[i][a];", "lRes[a][i] += wGp * hph * DN[i][a];"),
    ("vms subgrid convection sign", "lRes[a][i] += -wGp * sdu[g][i] * ahDN;", "lRes[a][i] += wGp * sdu[g][i] * ahDN;"),
    ("vms subgrid convection dropped", "lRes[a][i] += -wGp * sdu[g][i] * ahDN;", ""),
    ("vms continuity invEps dropped", "lRes[a][3] += wGp * trGradHdu * VLN[g][a] * invEpsilon;", "lRes[a][3] += wGp * trGradHdu * VLN[g][a];"),
    ("vms continuity subgrid sign", "lRes[a][3] += -wGp * sduhDN * invEpsilon;", "lRes[a][3] += wGp * sduhDN * invEpsilon;"),
    ("vms nodal mass V/3", "s->nd_mass[ids[i]] += s->vol[e] / 4.0;", "s->nd_mass[ids[i]] += s->vol[e] / 3.0;"),
    ("vms update sign", "s->sol_tau[i * 4 + j] = s->sol_n[i * 4 + j] - s->p.dual_step", "s->sol_tau[i * 4 + j] = s->sol_n[i * 4 + j] + s->p.dual_step"),
    ("vms subscale pressure-gradient sign", "sdu[g][i] / dt - ahGradDu[i] - gradP[i];", "sdu[g][i] / dt - ahGradDu[i] + gradP[i];"),
    ("vms subscale dt dropped", "sdu[g][i] / dt - ahGradDu[i] - gradP[i];", "sdu[g][i] - ahGradDu[i] - gradP[i];"),
    ("vms projection pressure sign", "projL[a][i] += wGp * (ahGradDu[i] + gradP[i]) * VLN[g][a];", "projL[a][i] += wGp * (ahGradDu[i] - gradP[i]) * VLN[g][a];"),
    ("vms subscale projection sign", "s->sdus[k] = (s->sdus[k] + projIGp) * tau1;", "s->sdus[k] = (s->sdus[k] - projIGp) * tau1;"),
    ("vms tau_u advective h^2", "c2 * maxA / insDiameter);", "c2 * maxA / (insDiameter * insDiameter));"),
    ("vms tau_p c1 dropped", "double tau_p = (insDiameter * insDiameter) / (c1 * tau_u);", "double tau_p = (insDiameter * insDiameter) / (tau_u);"),
    ("supg K1 rho dropped", "t1 += rho * L[c][A] * US[c][Bc] * detJ * w;", "t1 += L[c][A] * US[c][Bc] * detJ * w;"),
    ("supg convective gradient transposed", "for (int k = 0; k < 3; k++) U[c][c * 3 + k] = u_gp[k];", "for (int k = 0; k < 3; k++) U[c][k * 3 + c] = u_gp[k];"),
    ("supg shear row wrong", "B[3][a * 3 + 0] = d[1]; B[3][a * 3 + 1] = d[0];", "B[3][a * 3 + 0] = d[0]; B[3][a * 3 + 1] = d[1];"),
    ("supg D normal 2mu -> mu", "D[0][0] = lam + 2 * mu;", "D[0][0] = lam + mu;"),
    ("supg D shear 2mu", "D[3][3] = mu; D[4][4] = mu;", "D[3][3] = 2 * mu; D[4][4] = mu;"),
    ("supg K3 tau dropped", "t1 += rho * tau * US[c][A] * US[c][Bc] * detJ * w;", "t1 += rho * US[c][A] * US[c][Bc] * detJ * w;"),
    ("supg K4 sign", "K4[A][Bc] += rho * nuC * H[A] * H[Bc] * detJ * w;", "K4[A][Bc] -= rho * nuC * H[A] * H[Bc] * detJ * w;"),
    ("supg tau viscous nu^1", "t3 = t3 * (mu / rho) * (mu / rho) * cI;", "t3 = t3 * (mu / rho) * cI;"),
    ("supg tau cI dropped", "t3 = t3 * (mu / rho) * (mu / rho) * cI;", "t3 = t3 * (mu / rho) * (mu / rho);"),
    ("supg nuC trace dropped", "*nuC = 1.0 / (trG * *tau);", "*nuC = 1.0 / (*tau);"),
    ("supg lambda alpha dropped", "*lam = alpha * ((l1 < l2) ? l2 : l1);", "*lam = ((l1 < l2) ? l2 : l1);"),
    ("supg lambda G instead of G^-1", "for (int b = 0; b < 3; b++) iGu[a] += iG[a][b] * u[b];\n  }\n  double l2", "for (int b = 0; b < 3; b++) iGu[a] += Gm[a][b] * u[b];\n  }\n  double l2"),
    ("supg regime swapped", "*regime = (l1 > l2) ? 1.0 : 2.0;", "*regime = (l1 > l2) ? 2.0 : 1.0;"),
]
MUTANTS.append(("supg update mass inverted",
                "s->sol_n[i * 3 + k] = s->sol_p[i * 3 + k] + (1.0 / s->l_M[i * 3 + k]) * s->sol_n[i * 3 + k];",
                "s->sol_n[i * 3 + k] = s->sol_p[i * 3 + k] + (s->l_M[i * 3 + k]) * s->sol_n[i * 3 + k];"))


def _run_mutant(args):
    i, (name, old, new), src, tmp = args
    c = os.path.join(tmp, f"m{i}.c")
    so = os.path.join(tmp, f"m{i}.so")
    assert src.count(old) == 1, name
    open(c, "w").write(src.replace(old, new))
    b = subprocess.run(["make", "-s", "-C", ORACLE, "variant", f"VSRC={c}", f"VOUT={so}"], capture_output=True,
                       text=True)
    if b.returncode != 0:
        return name, "build failed: " + b.stderr[-400:]
    env = dict(os.environ, FEMX_ORACLE_LIB=so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *CHECKS],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=300)
    caught = r.returncode == 1 and " failed" in r.stdout
    return name, None if caught else "not caught:\n" + r.stdout[-600:]


def test_every_planted_mistake_is_caught(tmp_path):
    src = open(os.path.join(ORACLE, "femx_oracle.c")).read()
    jobs = [(i, m, src, str(tmp_path)) for i, m in enumerate(MUTANTS)]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(_run_mutant, jobs))
    missed = [(n, msg) for n, msg in res if msg]
    assert not missed, missed
    print(f"{len(res)} mutants, all caught")


def test_unmutated_copy_passes(tmp_path):
    """The same recipe on a copy whose only change is in a comment passes, so
    a caught mutant is the planted mistake and not the build."""
    src = open(os.path.join(ORACLE, "femx_oracle.c")).read()
    name, msg = _run_mutant((0, ("comment only", "TEST INFRASTRUCTURE ONLY", "TEST INFRASTRUCTURE ONLY (copy)"),
                             src, str(tmp_path)))
    assert msg and msg.startswith("not caught"), msg
