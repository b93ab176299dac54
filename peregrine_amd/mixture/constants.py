"""Physical constants in the units the mixture tables use (kmol, J, K).

Values follow ``peregrine/simulator/mixture/constants.py:1-14`` (CODATA 2018,
exact SI definitions).
"""

# J/kmol/K
Ru = 8314.46261815324
# J/K
kB = 1.380649e-23
# 1/kmol
avogadro = 6.02214076e26
# F/m
epsilon0 = 8.854187812773345e-12
# m/s, exact
clight = 299792458.0
# C m
debye = 1.0e-21 / clight
# K: the standard state of formation enthalpies and entropies
Tref = 298.15
# Pa: the standard pressure the equilibrium constants are referred to
pref = 101325.0
