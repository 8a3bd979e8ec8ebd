"""Initial-condition expressions: the reference's grammar and checks, and
array evaluation equal to scalar evaluation (Python floats and `math`, as
the reference evaluates one cell at a time) bit for bit."""

import math

import numpy as np
import pytest

from meshcfd_amd.expressions import ExpressionError, compile_expression, evaluate_initial_conditions
from meshcfd_amd.model import InitialConditionBlock, InitialConditionSettings, InitialState

FIELDS = [
    "1.225 + 0.1*sin(2*pi*x)*cos(pi*y)",
    "x^2 + y**3 - 2*x*y / (1 + x*x)",
    "exp(-((x-0.5)**2 + (y-0.5)**2) / 0.01) + 1",
    "max(x, y, 0.3) - min(abs(x - y), 0.2) + hypot(x, y) + atan2(y, x + 2)",
    "sqrt(x + 2) * log(3 + y) + log10(5) + floor(3*x) + ceil(2*y) + fabs(-x) + pow(2, x)",
    "101325 * (1 + 0.2 * tanh(10*(x - 0.5)))",
    "(x % 0.3) + degrees(y) / 360 + radians(30)",
]
DOMAINS = ["x < 0.5", "x < 0.5 and y >= 0.25", "not (x > 0.2) or y == y", "0.1 < x <= 0.7", "True",
           "(x - 0.5)^2 + (y - 0.5)^2 < 0.1^2"]


@pytest.mark.parametrize("source", FIELDS)
def test_field_arrays_equal_scalar_evaluation(source):
    rng = np.random.default_rng(3)
    x, y = rng.random(200), rng.random(200)
    e = compile_expression(source)
    arr = e.evaluate(x, y)
    ref = np.array([e.evaluate(float(a), float(b)) for a, b in zip(x, y)])
    assert np.array_equal(arr, ref)
    # and the scalar path is plain Python arithmetic with math
    if source == FIELDS[0]:
        a, b = 0.37, 0.81
        assert e.evaluate(a, b) == 1.225 + 0.1 * math.sin(2 * math.pi * a) * math.cos(math.pi * b)


@pytest.mark.parametrize("source", DOMAINS)
def test_domain_arrays_equal_scalar_evaluation(source):
    rng = np.random.default_rng(4)
    x, y = rng.random(300), rng.random(300)
    e = compile_expression(source, boolean=True)
    arr = e.evaluate(x, y)
    ref = np.array([e.evaluate(float(a), float(b)) for a, b in zip(x, y)])
    assert arr.dtype == bool and np.array_equal(arr, ref)


@pytest.mark.parametrize("source, boolean, message", [
    ("", False, "must not be empty"),
    ("x +", False, "Invalid expression"),
    ("__import__('os')", False, "Unsupported function call"),
    ("x.real", False, "Unsupported expression syntax"),
    ("z + 1", False, "Unknown expression name 'z'"),
    ("x < 1", False, "Field expression must be numeric"),
    ("x + 1", True, "Domain expression must be a boolean predicate"),
    ("2 ** 2000", False, "Exponent magnitude"),
    ("1" + "+1" * 100, False, "too complex"),
    ("x" * 600, False, "too long"),
])
def test_invalid_expressions_are_rejected(source, boolean, message):
    with pytest.raises(ExpressionError, match=message):
        compile_expression(source, boolean=boolean).evaluate(0.5, 0.5)


def test_field_must_be_finite():
    with pytest.raises(ExpressionError, match="finite"):
        compile_expression("1 / (x - 0.5)").evaluate(np.array([0.5, 0.2]), np.array([0.0, 0.0]))
    with pytest.raises(ExpressionError):
        compile_expression("1 / (x - 0.5)").evaluate(0.5, 0.0)


def test_blocks_apply_the_last_matching_block():
    settings = InitialConditionSettings(
        base_state=InitialState(1.0, 0.0, 0.0, 1.0),
        blocks=[
            InitialConditionBlock("left", "x < 0.5", "2", "x", "0", "3"),
            InitialConditionBlock("corner", "x < 0.25 and y < 0.25", "4", "0", "y", "5"),
        ],
    )
    c = np.array([[0.1, 0.1], [0.3, 0.1], [0.8, 0.9]])
    out = evaluate_initial_conditions(c, settings)
    assert out[0, :4].tolist() == [4.0, 0.0, 0.1, 5.0]
    assert out[1, :4].tolist() == [2.0, 0.3, 0.0, 3.0]
    assert out[2, :4].tolist() == [1.0, 0.0, 0.0, 1.0]
    assert out[:, 4:].tolist() == [[1e-6, 1.0]] * 3
    bad = InitialConditionSettings(blocks=[InitialConditionBlock("neg", "x > 0.5", "-1", "0", "0", "1")])
    with pytest.raises(ExpressionError, match="Initial-condition block 'neg', cell 2"):
        evaluate_initial_conditions(c, bad)
