"""Physical constants with SIMsalabim 5.36's values (reference: src/jaxdd/constants.py:8-12).

They are not CODATA's. The solution is compared with SIMsalabim's to 1e-5 of the current scale,
and a constant that differs in its fifth digit moves the thermal voltage and every exponential
with it, so these are kept bit-equal to the reference's.
"""

#: Elementary charge (C).
Q = 1.6022e-19
#: Boltzmann's constant (J/K).
K_B = 1.3807e-23
#: Vacuum permittivity (F/m).
EPS_0 = 8.8542e-12
