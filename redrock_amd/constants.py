"""Constants of the chi2-scan path (values of py/redrock/constants.py:7-37)."""

max_velo_diff = 1000.0       # km/s, constants.py:7
max_velo_diff_star = 100.0   # km/s, constants.py:8
min_resolution_integral = 0.99   # constants.py:9
min_deltachi2 = 9.           # constants.py:11

LyA_wavelength = 1215.67     # constants.py:15

# Lyman-series optical-depth tables (constants.py:16-37).  Each model is an
# ordered dict of lines with tau_eff = A (lambda_obs/line)^B blueward of the
# line; the order is the order in which the factors are multiplied.
Lyman_series = dict()
Lyman_series['Calura12'] = {
    'Lya': {'line': LyA_wavelength, 'A': 0.0023, 'B': 3.64, 'var_evol': 3.8},
    'Lyb': {'line': 1025.72, 'A': 0.0023 / 5.2615, 'B': 3.64, 'var_evol': 3.8},
    'Ly3': {'line': 972.537, 'A': 0.0023 / 14.356, 'B': 3.64, 'var_evol': 3.8},
    'Ly4': {'line': 949.7431, 'A': 0.0023 / 29.85984, 'B': 3.64, 'var_evol': 3.8},
    'Ly5': {'line': 937.8035, 'A': 0.0023 / 53.36202, 'B': 3.64, 'var_evol': 3.8},
}
Lyman_series['Kamble20'] = {
    'Lya': {'line': LyA_wavelength, 'A': 0.00554, 'B': 3.182},
}

#- speed of light in km/s (scipy.constants.speed_of_light/1000, fitz.py:40)
c_kms = 299792.458
