"""Parameter objects for the drop-in solver functions.

The reference builds these from argparse (class_art_param.py:14-138,
class_sir_param.py:10-118, class_admm_param.py:14-182).  The solver functions read only
the attributes below; defaults are the reference's.
"""


class param:
    """Attribute bag with the reference defaults used on the SIRT/MLEM/SPS/CG/ADMM path."""

    DEFAULTS = dict(
        npix_op=None, nang=None, nz=1, n_iter=None, eps=None, pc=-1.0, tv=0.0,
        mask=None, mask_add=None, edge_padding=0.0, index_start=0, index_end=None,
        plot=False, checkit=False, projector="dist-driv", algorithm="sirt",
        num_cores=-1, reg=None, reg_cost=0.1, huber_cost=0.1, n_iter_pcg=4, lambd1=1.0,
        lambd2=1.0,
        mu=1.0, init_object=False, radon_degree=0,     # reg_cost / huber_cost: sir.py:128,145
    )

    def __init__(self, **kw):
        for k, v in self.DEFAULTS.items():
            setattr(self, k, v)
        for k, v in kw.items():
            setattr(self, k, v)


# the reference's three parameter classes carry the same attributes for this path
art_param = sir_param = admm_param = param

PROJECTOR_OPER = {"ray-driv": "rd", "dist-driv": "dd", "pix-driv": "pd", "slant": "ss"}
