"""WGSL field-body translation (CPU: translator + hiprtc compile, no device needed)."""
from __future__ import annotations

import re
from pathlib import Path

import pytest

import linon_amd as L
from linon_amd import configs

REF_FUNCTIONS = Path("/root/reference/src/functions.rs")


def tr(body: str) -> str:
    return L.wgsl_translate(body)


def err(body: str) -> str:
    with pytest.raises(L.LinonError) as e:
        L.wgsl_translate(body)
    assert e.value.status == 2
    return e.value.message


def test_float_literals_become_f32():
    out = tr("let a = 1.5; let b = .25; let c = 2.; let d = 1e-3; let e = 0x1p4; return v * a;")
    assert "1.5f" in out and "0.25f" in out and "2.0f" in out and "1e-3f" in out and "0x1p4f" in out
    assert "double" not in out


def test_abstract_int_takes_the_context_type():
    out = tr("let x = max(t, 0); let y = 2 * p; let n = 3; return y * x + v * f32(n);")
    assert "lw::max(t, 0.0f)" in out
    assert "(2.0f * p)" in out
    assert "const int n = 3" in out
    out = tr("var k: f32 = 1; return v * k;")
    assert "float k = 1.0f" in out


def test_wgsl_clamp_argument_order_is_kept():
    # linon calls clamp(0.0, 1.0, x), i.e. e = 0, low = 1, high = x
    assert "lw::clamp(0.0f, 1.0f, t)" in tr("let a = clamp(0.0, 1.0, t); return v * a;")


def test_control_flow_and_swizzles():
    out = tr("""
        var acc = vec3<f32>(0.0);
        for (var i = 0; i < 4; i++) {
            if i == 2 { continue; }
            acc += p.zyx * f32(i);
        }
        var k = 0;
        while (k < 2) { k += 1; }
        loop { break; }
        if (t > 1.0) { return acc; } else if (t > 0.5) { return acc.xyz; } else { return v; }
    """)
    assert "lw::swizzle3<2,1,0>(p)" in out
    assert "for (int i = 0; (i < 4); i++)" in out
    assert "while (true)" in out


def test_reserved_names_are_renamed():
    out = tr("let float = 1.0; let new = 2.0; return v * float * new;")
    assert "float_w" in out and "new_w" in out


@pytest.mark.parametrize("body,msg", [
    ("return 1.0;", "must return a vec3"),
    ("return;", "must return a vec3"),
    ("let x = ;", "unexpected ';'"),
    ("return v.w;", "unsupported component"),
    ("return foo(v);", "unknown function 'foo'"),
    ("return q;", "unknown identifier 'q'"),
    ("let a = vec2<f32>(1.0, 2.0); return v;", "vec2"),
    ("if (t) { return v; } return v;", "condition must be bool"),
    ("return v + ;", "unexpected"),
    ("let a = 1.0 return v;", "expected ';'"),
    ("return v; /* open", "unterminated"),
    ("let a = length(1.0, 2.0); return v;", "takes 1 argument"),
    ("return rotateY(1.0, v);", "vec3<f32> is required"),
])
def test_errors_have_line_and_column(body, msg):
    m = err(body)
    assert re.match(r"field_function:\d+:\d+: ", m), m
    assert msg in m, m


def test_error_line_numbers_count_body_lines():
    m = err("let a = 1.0;\nlet b = 2.0;\nreturn a;")
    assert m.startswith("field_function:3:")


@pytest.mark.parametrize("name", list(configs.FIELD_WGSL))
def test_benchmark_field_bodies_translate(name):
    out = tr(configs.FIELD_WGSL[name])
    assert "return" in out


def test_hiprtc_compiles_a_body_for_gfx950():
    assert "gfx950 code object" in L.wgsl_compile_check(configs.FIELD_WGSL["hot_air"])


def test_missing_return_on_some_path_is_rejected():
    m = err("let x = 1.0;\nif (x > 0.0) { return v; }")
    assert "must end with a return" in m
    tr("if (t > 0.0) { return v; } else { return v0; }")   # both paths return


def test_hiprtc_reports_errors_in_body_coordinates():
    # translates (the vocabulary call exists) but fails overload resolution in hiprtc
    with pytest.raises(L.LinonError) as e:
        L.wgsl_compile_check("let x = 1.0;\nreturn v * select(1u, 2u, t > x);")
    assert e.value.status == 2
    assert "field_function:2" in e.value.message


@pytest.mark.skipif(not REF_FUNCTIONS.exists(), reason="reference checkout not mounted")
def test_reference_predefined_bodies_translate_and_compile():
    """Every PredefinedFunction body of src/functions.rs, read from the reference checkout at
    test time, goes through the translator; two of them also through hiprtc."""
    text = REF_FUNCTIONS.read_text()
    text = text[text.index("fn to_code"):]
    bodies = [b for b in re.findall(r'=> "(.*?)"\s*\.to_string\(\)', text, flags=re.S) if b]
    assert len(bodies) == 9
    for b in bodies:
        assert "return" in tr(b)
    for b in (bodies[2], bodies[7]):
        L.wgsl_compile_check(b)


def test_integer_division_is_wgsls_total_division():
    """WGSL integer / and % never trap: x / 0 = x, INT_MIN / -1 = INT_MIN, and both remainders 0
    (C++ leaves those cases undefined), so integer division goes through lw::wdiv / lw::wmod."""
    out = tr("let a: i32 = i32(t); let b = 7 / a; let c: u32 = u32(t); let d = 5u / c; let e = a % 3; "
             "let g = 2.0 / t; return v * f32(b) * f32(d) * f32(e) * g;")
    assert "lw::wdiv(7, a)" in out and "lw::wdiv(5u, c)" in out and "lw::wmod(a, 3)" in out
    assert "(2.0f / t)" in out
