"""Shared pieces of the command-line drivers (art.py, sir.py, admm.py of this package):
input/output paths, single-file and stack input, sinogram pre-processing, projection
angles and projector selection, as the reference's drivers do them.

  get_io_path             utils.py:41-58
  read_image/write_image  my_image_io.py:52-80 (readImage / writeImage dispatch by
                          extension: .DMP and grey-scale .tif)
  read_sinograms          art.py:612-659 (one file with -i, else every .DMP / .tif of the
                          input directory, sorted)
  sino_edge_padding       my_image_process.py:140-164
  sino_correct_rot_axis   my_image_process.py:87-118
  create_projection_angles utils.py:130-158 (equally spaced on [0, 180) or a text file)
  make_projector          art.py:387-400 (-pr names)
"""
import glob
import os
import shutil

import numpy as np

from .dmp import read_dmp, write_dmp
from .tiff import read_tiff, write_tiff

myfloat = np.float32


def read_image(path):
    ext = os.path.splitext(path)[1].lower()
    if ext == ".dmp":
        return read_dmp(path)
    if ext in (".tif", ".tiff"):
        return read_tiff(path)
    raise ValueError("%s: only .DMP and .tif images are read" % path)


def write_image(path, image):
    ext = os.path.splitext(path)[1].lower()
    if ext == ".dmp":
        return write_dmp(path, image)
    if ext in (".tif", ".tiff"):
        return write_tiff(path, image)
    raise ValueError("%s: only .DMP and .tif images are written" % path)


def get_io_path(args):
    """utils.py:41-58: the output directory defaults to the input one; an existing output
    directory is emptied when a whole stack is reconstructed."""
    pathin = args.pathin if args.pathin.endswith("/") else args.pathin + "/"
    if args.pathout is None:
        return pathin, pathin
    pathout = args.pathout if args.pathout.endswith("/") else args.pathout + "/"
    if os.path.exists(pathout) and args.filein is None:
        shutil.rmtree(pathout)
        os.makedirs(pathout)
    elif not os.path.exists(pathout):
        os.makedirs(pathout)
    return pathin, pathout


def read_sinograms(args, pathin):
    """(list of float32 sinograms, list of their file names)."""
    if args.filein is not None:
        sino = read_image(pathin + args.filein).astype(myfloat)
        if getattr(args, "angle_pi", False):
            sino = sino[: sino.shape[0] - 1, :]
        return [sino], [args.filein]
    names = sorted(os.path.basename(f) for f in glob.glob(pathin + "*.DMP"))
    if not names:
        names = sorted(os.path.basename(f) for f in glob.glob(pathin + "*.tif"))
    if not names:
        raise SystemExit("\nERROR: no .DMP or .tif file found in:\n" + pathin)
    return [read_image(pathin + f).astype(myfloat) for f in names], names


def sino_edge_padding(sino, padding):
    """Repeat the first and last detector columns int(padding * npix) times on each side."""
    nang, npix = sino.shape
    npad = int(padding * npix)
    left = np.ones((1, npad)) * sino[:, 0].reshape(nang, 1)
    right = np.ones((1, npad)) * sino[:, npix - 1].reshape(nang, 1)
    return np.concatenate((left, sino, right), axis=1)


def sino_correct_rot_axis(sino, ctr):
    """Shift the sinogram so that the rotation axis at detector position ctr lands on
    npix / 2: an integer shift repeats the edge column, a fractional one interpolates
    (cubic, over a 100-pixel zero border, as the reference)."""
    nang, npix = sino.shape
    shift = ctr - npix * 0.5
    if not float(shift).is_integer():
        from scipy import interpolate
        npad = npix + 200
        pad = np.zeros((nang, npad), dtype=myfloat)
        i1 = int((npad - npix) * 0.5)
        i2 = i1 + npix
        pad[:, i1:i2] = sino
        gx1, gy1 = np.meshgrid(np.arange(npad), np.arange(nang))
        z = np.dstack((gy1, gx1)).reshape(npad * nang, 2)
        gx2, gy2 = np.mgrid[0:nang, i1 + shift:i2 + shift]
        return interpolate.griddata(z, pad.reshape(-1), (gx2, gy2), method="cubic")
    shift = int(ctr - npix / 2)
    out = np.roll(sino, -shift, axis=1)
    if shift > 0:
        block = np.ones((1, shift)) * np.array(sino[:, npix - 1]).reshape(nang, 1)
        return np.concatenate((out[:, :npix - shift], block), axis=1)
    block = np.ones((1, -shift)) * np.array(sino[:, 0]).reshape(nang, 1)
    return np.concatenate((block, out[:, -shift:]), axis=1)


def create_projection_angles(nang=None, textfile=None):
    """Degrees: nang equally spaced on [0, 180), or the tab-separated list of a file."""
    if textfile is None:
        return np.linspace(0.0, 180.0, nang, endpoint=False)
    return np.fromfile(textfile, sep="\t")


def projector_suffix(name):
    return {"grid-kb": "_grid_kb", "grid-pswf": "_grid_pswf", "bspline": "_bspline",
            "slant": "_slant", "pix-driv": "_pd", "ray-driv": "_rd",
            "dist-driv": "_dd"}.get(name, "")


def make_projector(name, npix, angles_deg, radon_degree=0):
    """art.py:387-400: the projector object of a -pr name."""
    if name == "grid-pswf":
        from .class_projectors_grid import projectors as grid
        return grid(npix, angles_deg, kernel="pswf", oversampl=2.0, radon_degree=radon_degree)
    if name == "grid-kb":
        from .class_projectors_grid import projectors as grid
        return grid(npix, angles_deg, kernel="kb", oversampl=1.5, W=6.6, errs=6.0e-6,
                    interp="lin", radon_degree=radon_degree)
    from .class_projectors_radon import projectors as radon
    from .params import PROJECTOR_OPER
    if name not in PROJECTOR_OPER:
        raise SystemExit("\nERROR: projector %r is not available (have grid-pswf, grid-kb, "
                         "slant, pix-driv, ray-driv, dist-driv)\n" % name)
    if radon_degree:
        raise SystemExit("\nERROR: selected projector %s cannot be used for differentiated "
                         "tomography!!\n" % name)
    return radon(npix, angles_deg, oper=PROJECTOR_OPER[name])


def add_common_args(parser, projector_default):
    """The options the three drivers share (art.py:91-200, sir.py:75-170, admm.py:104-212)."""
    parser.add_argument("-Di", "--pathin", dest="pathin", default="./",
                        help="Specify path to input data")
    parser.add_argument("-i", "--filein", dest="filein", help="Specify name of input sinogram")
    parser.add_argument("-Do", "--pathout", dest="pathout", help="Specify path to output data")
    parser.add_argument("-g", "--geometry", dest="geometry", default="0",
                        help="-g 0: equiangular on [0, 180); -g angles.txt: angles (degrees)")
    parser.add_argument("-c", dest="ctr", default=-1, type=myfloat, help="Centre of rotation")
    parser.add_argument("-z", dest="edge_padding", type=myfloat, default=0.0,
                        help="Enable edge padding of the sinogram")
    parser.add_argument("-lt", dest="lt", action="store_true",
                        help="Enable local tomography configuration")
    parser.add_argument("-pr", dest="projector", default=projector_default,
                        help="grid-pswf, grid-kb, slant, pix-driv, ray-driv, dist-driv")
    parser.add_argument("-p", dest="plot", action="store_true", help="(no plotting here)")
    parser.add_argument("-l", dest="logfile", action="store_true", help="Write log file")
    parser.add_argument("-init", dest="init_object", action="store_true",
                        help="Initialize reconstruction with FBP reconstruction")
    parser.add_argument("-cit", dest="checkit",
                        help="Save every iterate into <checkit>_rmse_folder/")
    parser.add_argument("-eps", dest="eps", type=myfloat, help="Select stopping epsilon")
    parser.add_argument("-pi", dest="angle_pi", action="store_true",
                        help="Eliminate last projection")
    parser.add_argument("-nc", dest="num_cores", type=int, default=-1,
                        help="accepted for compatibility; one GPU runs every slice")
    parser.add_argument("-resume", dest="resume",
                        help="start from this saved iterate (.DMP/.tif, or a directory of "
                             "them for a stack) instead of the usual start")
    parser.add_argument("-it0", dest="it0", type=int, default=0,
                        help="iterations the -resume iterate has already done")


def prepare(args):
    """Steps of art.py:617-760 shared by the drivers: read, edge padding, centre
    correction, angles.  Returns a dict."""
    pathin, pathout = get_io_path(args)
    sinos, names = read_sinograms(args, pathin)
    nang, npix = sinos[0].shape
    ctr = npix * 0.5 if args.ctr == -1 else float(args.ctr)
    edf = 0.87 if args.lt else (args.edge_padding if args.edge_padding != 0 else 0.0)
    npix_old = npix
    if edf:
        sinos = [sino_edge_padding(s, edf) for s in sinos]
        npix = sinos[0].shape[1]
        i1 = int((npix - npix_old) * 0.5)
        ctr += i1
    else:
        i1 = 0
    i2 = i1 + npix_old
    if args.ctr != -1:
        sinos = [sino_correct_rot_axis(s, ctr) for s in sinos]
    if args.geometry == "0":
        angles = create_projection_angles(nang)
    else:
        angles = create_projection_angles(textfile=pathin + args.geometry)
    return {"pathin": pathin, "pathout": pathout, "names": names,
            "sinos": np.stack(sinos).astype(myfloat), "npix": npix, "npix_old": npix_old,
            "nang": nang, "i1": i1, "i2": i2, "edf": edf, "angles": angles, "ctr": ctr}


def load_resume(args, nz, npix):
    """The -resume iterate(s) as an (nz, npix, npix) float32 array."""
    if os.path.isdir(args.resume):
        files = sorted(glob.glob(os.path.join(args.resume, "*.DMP"))) or \
            sorted(glob.glob(os.path.join(args.resume, "*.tif")))
    else:
        files = [args.resume]
    x = np.stack([read_image(f) for f in files]).astype(myfloat)
    if x.shape != (nz, npix, npix):
        raise SystemExit("\nERROR: -resume holds %s, the problem needs (%d, %d, %d)\n"
                         % (x.shape, nz, npix, npix))
    return x


def checkpoint_writer(args, root_label):
    """art.py:328-341: with -cit, the iterate of every step of the middle slice is written
    to <checkit>_rmse_folder/reco_iterNNN.DMP (the folder is emptied first), cropped to
    the reconstruction square.  Returns a callback (k, x, i1, i2) or None."""
    if args.checkit is None:
        return None
    path = args.checkit + "_rmse_folder/"
    if os.path.exists(path):
        shutil.rmtree(path)
    os.makedirs(path)

    def write(k, x, i1, i2):
        nz = x.shape[0]
        nzz = 0 if nz == 1 else int(0.5 * nz)
        write_dmp(path + "reco_iter%03d.DMP" % (k - 1), x[nzz, i1:i2, i1:i2])
    return write


def save_reco(pathout, filein, reco_name, root, reco):
    """art.py:479-493: -o names the output, else the input name with the method's root."""
    if reco_name is not None:
        filename = pathout + reco_name
    else:
        base, ext = filein[:-4], filein[-4:]
        filename = pathout + base + root + ext
    write_image(filename, reco)
    print("\nReconstruction saved in:\n", filename)
    return filename
