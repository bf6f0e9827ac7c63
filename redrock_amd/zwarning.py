"""ZWARN bit definitions (py/redrock/zwarning.py:14-56)."""


class ZWarningMask(object):
    SKY = 2**0
    LITTLE_COVERAGE = 2**1
    SMALL_DELTA_CHI2 = 2**2
    NEGATIVE_MODEL = 2**3
    MANY_OUTLIERS = 2**4
    Z_FITLIMIT = 2**5
    NEGATIVE_EMISSION = 2**6
    UNPLUGGED = 2**7
    BAD_TARGET = 2**8
    NODATA = 2**9
    BAD_MINFIT = 2**10
    POORDATA = 2**11
    RESERVED16 = 2**16
    RESERVED17 = 2**17
    RESERVED18 = 2**18
    RESERVED19 = 2**19
    RESERVED20 = 2**20
    RESERVED21 = 2**21
    RESERVED22 = 2**22
    RESERVED23 = 2**23

    @classmethod
    def flags(cls):
        import numpy as np
        flagmask = [(k, v) for k, v in cls.__dict__.items()
                    if not k.startswith('_') and k.isupper()]
        isort = np.argsort([x[1] for x in flagmask])
        return [flagmask[i] for i in isort]


#- zwarn bits that mark a bad individual template fit (zwarning.py:52-56)
badfit_mask = (ZWarningMask.NEGATIVE_MODEL | ZWarningMask.MANY_OUTLIERS
               | ZWarningMask.Z_FITLIMIT | ZWarningMask.NEGATIVE_EMISSION
               | ZWarningMask.BAD_MINFIT)
