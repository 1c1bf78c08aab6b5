"""Writes be_roots.json: roots of the backward-Euler radius/temperature
equations of particle_update_BE (particles.f90:3103-3178, ie_vrt_nd
4263-4327), solved at 50 significant digits with mpmath.

The equations are written here in their dimensional form, from the physics
(Fuchs/Maxwell diffusional growth with Ranz-Marshall ventilation, Koehler
curvature and solute terms, latent and sensible heat):

  G1 = r_n - r  - dt * rdot(r_n, T_n) = 0
  G2 = T_n - Tp - dt * Tdot(r_n, T_n) = 0

  rdot = Sh/(9 Sc) * (rho_p/rho_w) * (r_n/tau) * (q_inf - q*)
  Tdot = -Nu/(3 Pr) * (Cpa/Cpp) * (rho_p/rho_w) * (T_n - T_f)/tau + 3 Lv rdot/(r_n Cpp)
  q*   = Mw e_s(T_f)/(Ru T_n rho_a) * exp(Lv Mw/Ru (1/T_f - 1/T_n)
                                        + 2 Mw Gam/(Ru rho_w r_n T_n)
                                        - kappa m_s/(rho_s V(r_n)))
  tau  = rho_p (2 r_n)^2/(18 rho_a nu),  rho_p = rho_w + m_s/V(r_n),
  Re   = 2 r_n |u_f - v|/nu,  Sh = 2 + 0.6 Re^(1/2) Sc^(1/3),
  Nu   = 2 + 0.6 Re^(1/2) Pr^(1/3)

The code solves the same system non-dimensionalised (r_n = tr*r,
T_n = tt*Tp, time scaled by taup0); taup0 cancels, so these roots are the
code's roots.  Run: python tests/golden/make_be_roots.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50

C = dict(nuf=1.57e-5, Cpa=1006.0, Cpp=4179.0, Pra=0.715, Sc=0.615, rhow=1000.0, rhos=2000.0,
         Mw=0.018015, Ru=8.3144, Gam=7.28e-2, Lv=(25.0 - 0.02274 * 26.0) * 10.0 ** 5, rhoa=1.2)


def e_s(T):
    return mp.mpf("610.94") * mp.exp((mp.mpf("17.6257") * (T - mp.mpf("273.15"))) / (mp.mpf("243.04") + (T - mp.mpf("273.15"))))


def residual(rn, Tn, s):
    c = {k: mp.mpf(v) for k, v in C.items()}
    r, Tp, Tf, q, ms, kap, dt, du = (mp.mpf(s[k]) for k in ("r", "Tp", "Tf", "qinf", "m_s", "kappa", "dt", "slip"))
    V = mp.mpf(4) / 3 * mp.pi * rn ** 3
    rhop = c["rhow"] + ms / V
    tau = rhop * (2 * rn) ** 2 / (18 * c["rhoa"] * c["nuf"])
    Re = 2 * rn * du / c["nuf"]
    Sh = 2 + mp.mpf("0.6") * mp.sqrt(Re) * mp.cbrt(c["Sc"])
    Nu = 2 + mp.mpf("0.6") * mp.sqrt(Re) * mp.cbrt(c["Pra"])
    qs = c["Mw"] * e_s(Tf) / (c["Ru"] * Tn * c["rhoa"]) * mp.exp(
        c["Lv"] * c["Mw"] / c["Ru"] * (1 / Tf - 1 / Tn) + 2 * c["Mw"] * c["Gam"] / (c["Ru"] * c["rhow"] * rn * Tn)
        - kap * ms / (c["rhos"] * V))
    rdot = Sh / (9 * c["Sc"]) * (rhop / c["rhow"]) * (rn / tau) * (q - qs)
    Tdot = -Nu / (3 * c["Pra"]) * (c["Cpa"] / c["Cpp"]) * (rhop / c["rhow"]) * (Tn - Tf) / tau + 3 * c["Lv"] * rdot / (rn * c["Cpp"])
    return rn - r - dt * rdot, Tn - Tp - dt * Tdot


def q_of_rh(rh, T):
    c = {k: mp.mpf(v) for k, v in C.items()}
    return rh * c["Mw"] / c["Ru"] / T * e_s(T) / c["rhoa"]


def cases():
    out = []
    for r in (1.0e-6, 1.4e-6, 2.0e-6, 5.0e-6, 1.2e-5, 2.0e-5):
        for rh in (0.998, 1.0, 1.002):
            for dt in (1.0e-4, 5.0e-4, 5.0e-3):
                Tf = 300.0
                s = dict(r=r, Tp=Tf + 0.03, Tf=Tf, qinf=float(q_of_rh(mp.mpf(rh), mp.mpf(Tf))), m_s=5e-18,
                         kappa=1.2, dt=dt, slip=0.01)
                out.append(s)
    for ms in (1e-19, 1e-17, 1e-15):
        s = dict(r=1e-5, Tp=299.9, Tf=300.0, qinf=float(q_of_rh(mp.mpf("1.001"), mp.mpf(300))), m_s=ms, kappa=1.2,
                 dt=1e-3, slip=0.05)
        out.append(s)
    return out


def main():
    rows = []
    for s in cases():
        f = lambda rn, Tn: residual(rn, Tn, s)
        sol = None
        for start in ((s["r"], s["Tf"]), (s["r"] * 1.01, s["Tf"]), (s["r"] * 0.9, s["Tf"])):
            try:
                sol = mp.findroot(f, (mp.mpf(start[0]), mp.mpf(start[1])), tol=mp.mpf(10) ** -40, maxsteps=200)
                break
            except (ZeroDivisionError, ValueError):
                continue
        if sol is None:
            continue
        rn, Tn = sol[0], sol[1]
        g1, g2 = residual(rn, Tn, s)
        assert abs(g1) < 1e-35 and abs(g2) < 1e-30, (g1, g2)
        rows.append(dict(s, r_new=mp.nstr(rn, 25), Tp_new=mp.nstr(Tn, 25)))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "be_roots.json")
    with open(path, "w") as fh:
        json.dump({"constants": C, "roots": rows}, fh, indent=1)
    print(f"wrote {len(rows)} roots to {path}")


if __name__ == "__main__":
    main()
