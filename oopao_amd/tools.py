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


def strehlMeter(PSF, tel, PSF_ref=None, zeroPaddingFactor=2, display=False, title=''):
    """Strehl ratio [%] from the OTF of a focal-plane image (tools/tools.py:102-157), on the GPU.
    Without PSF_ref the diffraction-limited image is computed after resetting the OPD, as the
    reference does (tel.resetOPD(); tel.computePSF(zeroPaddingFactor))."""
    import torch
    from .science import strehl_dev
    if PSF_ref is None:
        tel.resetOPD()
        tel.computePSF(zeroPaddingFactor)
        ref = tel._PSF
    else:
        ref = torch.as_tensor(np.asarray(PSF_ref) if not isinstance(PSF_ref, torch.Tensor) else PSF_ref,
                              device=tel.device)
    psf = PSF if isinstance(PSF, torch.Tensor) else torch.as_tensor(np.asarray(PSF), device=tel.device)
    sr = strehl_dev(psf.to(tel.device), ref.to(tel.device), tel.resolution, zeroPaddingFactor)
    if display:
        print('Strehl ratio [%] : ', sr)
    return sr
