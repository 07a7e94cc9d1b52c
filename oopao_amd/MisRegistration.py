"""Mis-registration parameters of a DM (MisRegistration.py:11-60): all zero by default."""


class MisRegistration:
    def __init__(self, param=None):
        self.tag = 'misRegistration'
        self.rotationAngle = 0
        self.shiftX = 0
        self.shiftY = 0
        self.anamorphosisAngle = 0
        self.tangentialScaling = 0
        self.radialScaling = 0
        if param is not None:
            for k in ('rotationAngle', 'shiftX', 'shiftY', 'anamorphosisAngle', 'tangentialScaling',
                      'radialScaling'):
                if isinstance(param, dict):
                    if k in param:
                        setattr(self, k, param[k])
                elif hasattr(param, k):
                    setattr(self, k, getattr(param, k))
        self.prop = {}

    def is_zero(self):
        return (self.rotationAngle == 0 and self.shiftX == 0 and self.shiftY == 0 and
                self.anamorphosisAngle == 0 and self.tangentialScaling == 0 and self.radialScaling == 0)
