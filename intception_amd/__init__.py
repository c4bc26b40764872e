"""intception_amd: MI355X-native contracted Gaussian ERI evaluation.

Input-script vocabulary (as in the reference's ``from intception import *``,
src/intception/__init__.py:28-34) plus this build's HIP backend.
"""
from .dsl import (boys_f, boys_function, dsl_base, dsl_binop, dsl_cartesian_gaussian, dsl_function,  # noqa: F401
                  dsl_integer_index, dsl_integral, dsl_position, dsl_rr, dsl_scalar, dsl_unop, dsl_value,
                  dsl_zero, exp, nint, norm, sqrt)
from .generator import generator, generator_options  # noqa: F401

__all__ = ["boys_f", "boys_function", "dsl_rr", "dsl_integral", "dsl_base", "dsl_binop", "dsl_cartesian_gaussian",
           "dsl_function", "dsl_integer_index", "dsl_position", "dsl_scalar", "dsl_unop", "dsl_value", "dsl_zero",
           "sqrt", "exp", "norm", "nint", "generator", "generator_options"]
