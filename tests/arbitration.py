"""Loader of the arbitration fixtures (tests/golden/config2_arbitration.npz,
config4_arbitration.npz): quartets on which the FP64 GPU path and the oracle
differed beyond the FP64 bar, with selected elements and their 40-digit
McMurchie-Davidson values."""
import os

import numpy as np

from intception_amd import basis


def arbitration_batch(name):
    """(batch, fixture) for tests/golden/<name>_arbitration.npz; element k of
    the fixture is b.out_off[elem_q[k]] + elem[k] in the batch's output."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "%s_arbitration.npz" % name))
    nq = len(g["l"]) // 4
    qs = []
    for q in range(nq):
        qs.append([{"l": int(g["l"][4 * q + k]), "exps": g["exps"][4 * q + k][:g["nprim"][4 * q + k]],
                    "coefs": g["coefs"][4 * q + k][:g["nprim"][4 * q + k]], "cen": g["cen"][4 * q + k]}
                   for k in range(4)])
    return basis.batch_from_shells(qs, True), g
