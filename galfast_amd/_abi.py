"""ctypes mirror of the parameter blocks declared in include/galfast_hip.h."""
import ctypes as C

f32 = C.c_float
f64 = C.c_double
i32 = C.c_int32
u32 = C.c_uint32
u64 = C.c_uint64
vp = C.c_void_p


class gf_col(C.Structure):
    _fields_ = [("ptr", vp), ("pitch", u32), ("width", u32), ("elem", u32), ("reserved", u32)]


class gf_rows(C.Structure):
    _fields_ = [("begin", u32), ("step", u32), ("end", u32), ("shmem", u32)]


class gf_bitmap(C.Structure):
    _fields_ = [("bits", u32 * 2)]


def bitmap(*comps):
    b = gf_bitmap()
    for c in comps:
        b.bits[c // 32] |= 1 << (c % 32)
    return b


class gf_rng(C.Structure):
    _fields_ = [("state", vp), ("nstreams", u32), ("reserved", u32)]


class gf_FeH_data(C.Structure):
    _fields_ = [("A", f32 * 2), ("sigma", f32 * 3), ("offs", f32 * 3),
                ("Hmu", f32), ("muInf", f32), ("DeltaMu", f32),
                ("comp_thin", gf_bitmap), ("comp_thick", gf_bitmap), ("comp_halo", gf_bitmap),
                ("M", (f32 * 3) * 3), ("T", f32 * 3)]


class gf_GaussianFeH_data(C.Structure):
    _fields_ = [("ncomp", i32), ("mean", f32 * 10), ("sigma", f32 * 10), ("pcum", f32 * 10)]


class gf_vmoments(C.Structure):
    _fields_ = [("m", (f32 * 5) * 3), ("ss", (f32 * 5) * 6)]


class gf_Bond2010_data(C.Structure):
    _fields_ = [("comp_thin", gf_bitmap), ("comp_thick", gf_bitmap), ("comp_halo", gf_bitmap),
                ("fk", f32), ("disk_1", gf_vmoments), ("disk_2", gf_vmoments), ("halo", gf_vmoments),
                ("M", (f32 * 3) * 3), ("T", f32 * 3)]


class gf_vel2pm_data(C.Structure):
    _fields_ = [("coordsys", i32), ("vLSR", f32), ("u0", f32), ("v0", f32), ("w0", f32), ("Rg", f32)]


class gf_photometry_data(C.Structure):
    _fields_ = [("ncolors", u32), ("bidx", u32), ("reddening", f32 * 17),
                ("nFeH", i32), ("nMr", i32), ("tcFeH", f32 * 2), ("tcMr", f32 * 2),
                ("color", vp * 4), ("cflags", vp * 4), ("texels", vp), ("flagbits", vp)]


class gf_tex1d(C.Structure):
    _fields_ = [("data", vp), ("n", i32), ("tc", f32 * 2)]


class gf_multiples_data(C.Structure):
    _fields_ = [("secProb", gf_tex1d), ("cumLF", gf_tex1d), ("invCumLF", gf_tex1d), ("algo", i32)]


class gf_photoerr_data(C.Structure):
    _fields_ = [("nbands", i32), ("koff", i32 * 18), ("kx", vp), ("ky", vp), ("kq", vp),
                ("gx0", C.c_double * 18), ("gscale", C.c_double * 18)]


class gf_chain(C.Structure):
    _fields_ = [("feh", C.POINTER(gf_FeH_data)), ("gfeh", C.POINTER(gf_GaussianFeH_data)),
                ("gfeh_comps", gf_bitmap), ("mult", C.POINTER(gf_multiples_data)),
                ("mult_comps", gf_bitmap), ("bond", C.POINTER(gf_Bond2010_data)),
                ("phot", C.POINTER(gf_photometry_data)), ("phot_comps", gf_bitmap),
                ("perr", C.POINTER(gf_photoerr_data)), ("v2pm", C.POINTER(gf_vel2pm_data)),
                ("gal2other_coordsys", i32), ("pad", i32)]


TABLE_COLUMNS = ["lb", "projIdx", "projXY", "DM", "absmag", "XYZ", "Am", "AmInf", "comp", "hidden",
                 "FeH", "FeH_comp", "absmagSys", "absmagSysNcomp", "vcyl", "mags", "photoFlags",
                 "obsMags", "pm", "radec"]


class gf_table(C.Structure):
    _fields_ = [(n, gf_col) for n in TABLE_COLUMNS]


class gf_lambert(C.Structure):
    _fields_ = [("l0", f64), ("b0", f64), ("cl", f32), ("cb", f32), ("sl", f32), ("sb", f32)]


class gf_pencil_beam(C.Structure):
    _fields_ = [("cl", f32), ("cb", f32), ("sl", f32), ("sb", f32), ("coveredFraction", f32),
                ("dx", f32), ("dA", f32), ("projIdx", i32), ("X", f32), ("Y", f32),
                ("extIdx", i32), ("pad", i32)]


class gf_skygen_params(C.Structure):
    _fields_ = [("m0", f32), ("m1", f32), ("dm", f32), ("M0", f32), ("M1", f32), ("dM", f32),
                ("dmin", f32), ("dmax", f32), ("npixels", i32), ("nm", i32), ("nM", i32),
                ("nthreads", i32), ("stopstars", i32), ("pad", i32), ("proj", gf_lambert * 2)]


GF_MODEL_EXPDISK = 1
GF_MODEL_POWERLAW = 2


class gf_model(C.Structure):
    _fields_ = [("kind", i32), ("comp", i32), ("f", f32), ("l", f32), ("h", f32), ("r_cut2", f32),
                ("n", f32), ("ba", f32), ("ca", f32), ("rminSq", f32), ("rmaxSq", f32),
                ("T", f32 * 3), ("M", (f32 * 3) * 3), ("norm", f32), ("lf", gf_tex1d)]


class gf_tex3d(C.Structure):
    _fields_ = [("data", vp), ("nx", i32), ("ny", i32), ("nz", i32), ("tc", (f32 * 2) * 3)]


class gf_extinction(C.Structure):
    _fields_ = [("north", gf_tex3d), ("south", gf_tex3d), ("beam", gf_tex3d)]


GF_FOOT_MAX = 8


class gf_foot_region(C.Structure):
    _fields_ = [("kind", i32), ("equ", i32), ("l0", f64), ("l1", f64), ("b0", f64), ("b1", f64),
                ("cl", f64), ("cb", f64), ("sl", f64), ("sb", f64), ("cosr", f64), ("hole", f64),
                ("bmin", f64)]


class gf_footprint(C.Structure):
    _fields_ = [("nregion", i32), ("pad", i32), ("r", gf_foot_region * GF_FOOT_MAX)]


class gf_count_result(C.Structure):
    _fields_ = [("nstarsExpectedToGenerate", f64), ("nstarsExpected", f64), ("maxCount", f64),
                ("ncells", f64)]


SKYGEN_COLUMNS = ["lb", "projIdx", "projXY", "DM", "M", "XYZ", "Am", "AmInf", "comp", "hidden"]


class gf_skygen_cols(C.Structure):
    _fields_ = [(n, gf_col) for n in SKYGEN_COLUMNS]


class gf_countsmap_params(C.Structure):
    _fields_ = [("x0", f32), ("dx", f32), ("y0", f32), ("dy", f32), ("z0", f32), ("dz", f32),
                ("nx", i32), ("ny", i32), ("nz", i32), ("zmag_width", i32), ("ncoadd", i32), ("equalarea", i32),
                ("coadd_offsets", vp), ("proj", (f64 * 4) * 2)]


GF_ACC_FP64, GF_ACC_KAHAN32, GF_ACC_FP32 = 0, 1, 2
