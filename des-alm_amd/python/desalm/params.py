"""ctypes mirror of ``desalm_params`` (include/desalm.h).

The fields mirror the module state that the reference's mg_driver reads
(src/global.F90, src/global_param.F90, src/global_app.F90 as filled by
init_input_file, src/init.F90:176-228, and init_daf_bc, src/init.F90:486-508).
Indices are 0-based here where the reference's are 1-based (iter, btype, ep).
"""
import ctypes

MAX_LEVELS = 6

PREC_FP64 = 0
PREC_MIXED = 1
# local pseudo time step formula (DESALM_DTAU_*)
DTAU_LIVE = 0
DTAU_MET = 1

BC_INTERFACE = 0
BC_WALL = 1
BC_SYMMETRY = 2
BC_FREESTREAM = 3
BC_INFLOW = 4
BC_OUTLET_MOC = 5
BC_PERIODIC = 6
BC_PERIODIC_WRAP = 7
BC_FREE_SURFACE = 8


class Params(ctypes.Structure):
    _fields_ = [
        ("ng", ctypes.c_int32),
        ("img", ctypes.c_int32), ("jmg", ctypes.c_int32), ("kmg", ctypes.c_int32),
        ("icrs", ctypes.c_int32), ("jcrs", ctypes.c_int32), ("kcrs", ctypes.c_int32),
        ("iter", ctypes.c_int32 * MAX_LEVELS),
        ("itm", ctypes.c_int32),
        ("it_min", ctypes.c_int32),
        ("er_min", ctypes.c_double),
        ("eo_min", ctypes.c_double),
        ("ren", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("delti", ctypes.c_double),
        ("unsteady", ctypes.c_int32),
        ("des", ctypes.c_int32),
        ("itk", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("cfl1", ctypes.c_double), ("vnn1", ctypes.c_double),
        ("cfl2", ctypes.c_double), ("vnn2", ctypes.c_double),
        ("pdiss_coef", ctypes.c_double),
        ("ep", (ctypes.c_double * 3) * 2),
        ("btype", ctypes.c_int32 * 6),
        ("bdir", ctypes.c_int32 * 6),
        ("ifix", ctypes.c_int32), ("jfix", ctypes.c_int32), ("kfix", ctypes.c_int32),
        ("dtau_formula", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 6),
        ("body_force", ctypes.c_double * 3),
    ]

    def level_dims(self):
        """Node counts of every level (src/init.F90:254-260)."""
        dims = [(self.img, self.jmg, self.kmg)]
        for _ in range(1, self.ng):
            i, j, k = dims[-1]
            dims.append(((i + self.icrs) // 2 ** self.icrs,
                         (j + self.jcrs) // 2 ** self.jcrs,
                         (k + self.kcrs) // 2 ** self.kcrs))
        return dims

    def nodes_all(self):
        return sum(i * j * k for i, j, k in self.level_dims())

    def copy(self):
        p = Params()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(self), ctypes.sizeof(Params))
        return p


def make_params(img, jmg, kmg, ng=2, crs=(1, 1, 1), iters=(1, 1, 1, 1, 1, 1), itm=1,
                it_min=None, er_min=-8.0, eo_min=-30.0, ren=1000.0, beta=1.0, delti=0.01,
                unsteady=False, des=False, itk=2, precision=PREC_FP64,
                cfl1=1.0, vnn1=0.1, cfl2=0.5, vnn2=0.05, pdiss=0.1,
                ep=((1.0, 2.5, 2.5), (0.1, 0.2, 0.2)),
                btype=(1, 1, 1, 1, 1, 1), bdir=(1, 1, 2, 2, 3, 3), fix=(2, 2, 2), body_force=(0.0, 0.0, 0.0),
                dtau_formula=0):
    """Build a Params.  Defaults follow the shipped deck (ind3dmg.dat:8-16):
    cfl1=1.0, vnn1=0.1, pdiss=0.1, ep=(1.0,2.5,2.5; 0.1,0.2,0.2), itk=2,
    cfl2=0.5, vnn2=0.05.  it_min defaults to itm (no early exit)."""
    p = Params()
    p.ng = ng
    p.img, p.jmg, p.kmg = img, jmg, kmg
    p.icrs, p.jcrs, p.kcrs = crs
    for n in range(MAX_LEVELS):
        p.iter[n] = iters[n] if n < len(iters) else 1
    p.itm = itm
    p.it_min = itm if it_min is None else it_min
    p.er_min, p.eo_min = er_min, eo_min
    p.ren, p.beta, p.delti = ren, beta, delti
    p.unsteady = int(bool(unsteady))
    p.des = int(bool(des))
    p.itk = itk
    p.precision = precision
    p.cfl1, p.vnn1, p.cfl2, p.vnn2 = cfl1, vnn1, cfl2, vnn2
    p.pdiss_coef = pdiss
    for a in range(2):
        for d in range(3):
            p.ep[a][d] = ep[a][d]
    for b in range(6):
        p.btype[b] = btype[b]
        p.bdir[b] = bdir[b]
    p.ifix, p.jfix, p.kfix = fix
    p.dtau_formula = dtau_formula
    for c in range(3):
        p.body_force[c] = body_force[c]
    return p
