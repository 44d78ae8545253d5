import sys, math, numpy as np
sys.path.insert(0, '/root/repo'); sys.path.insert(0, '/root/repo/tests'); sys.path.insert(0, '/root/repo/_work/abl')
import fvabl as fv
from test_oracle import orszag_tang_reference_ic
s, box = orszag_tang_reference_ic()
gam=5/3
def tot(x):
    eg = x[3]/(gam-1)+0.5*x[0]*(x[1]**2+x[2]**2); em = 0.5*(x[4]**2+x[5]**2+x[6]**2)
    return math.fsum(np.concatenate([eg.ravel(), em.ravel()]))
tol = float(sys.argv[1]); nsteps = int(sys.argv[2])
oc = fv.make_config(s.shape, 2, box_size=box, C_cfl=0.4, t_end=3.0, gamma=5/3, tol=tol)
st = s.copy(); E0 = tot(st)
dA = dB = 0.0
for k in range(nsteps):
    dt = fv.cfl_dt(st[:4], oc)
    e0 = tot(st)
    st[:4] = fv.gas_rk2(st[:4], oc, dt/2); e1 = tot(st)
    B, gas, it = fv.magnetic_update(st[4:], st[:4], oc, dt); st[4:] = B; st[:4] = gas; e2 = tot(st)
    st[:4] = fv.gas_rk2(st[:4], oc, dt/2); e3 = tot(st)
    dA += (e1-e0)+(e3-e2); dB += e2-e1
    if k % 100 == 99: print(k+1, f"cum hydro {dA:+.3e} cum mag {dB:+.3e} total {tot(st)-E0:+.3e}", flush=True)
