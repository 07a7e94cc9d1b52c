"""Error type and helpers shared by the drop-in classes (tools/tools.py:464-483, 558-570)."""
import numpy as np


class OopaoError(Exception):
    """Exception raised for errors related to OOPAO objects (tools/tools.py:562-570)."""

    def __init__(self, string):
        self.string = string
        self.message = string
        super().__init__(self.message)


def warning(string):
    """ANSI warning print (tools/tools.py:558-559)."""
    print('\033[01m\033[33m' + 'OOPAO Warning: \n' + string + '\033[0m')


def set_binning(array, binning_factor, mode='sum'):
    """Sum/mean binning of a 2-D or 3-D numpy array (tools/tools.py:464-483)."""
    if array.shape[0] % binning_factor != 0:
        raise ValueError('Binning factor %d not compatible with the array size' % binning_factor)
    if array.ndim == 2:
        new = [int(np.round(array.shape[0] / binning_factor)), int(np.round(array.shape[1] / binning_factor))]
        shape = (new[0], array.shape[0] // new[0], new[1], array.shape[1] // new[1])
        r = array.reshape(shape)
        return r.sum(-1).sum(1) if mode == 'sum' else r.mean(-1).mean(1)
    new = [int(np.round(array.shape[0] / binning_factor)), int(np.round(array.shape[1] / binning_factor)),
           array.shape[2]]
    shape = (new[0], array.shape[0] // new[0], new[1], array.shape[1] // new[1], new[2])
    r = array.reshape(shape)
    return r.sum(-2).sum(1) if mode == 'sum' else r.mean(-2).mean(1)
