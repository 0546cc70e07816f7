"""Symmetry scripts on the host (no GPU): the library's Lua interpreter and the
rows of LuaSymmetry::PreCalculate (Backend/Backend/Symmetry.cpp:463-521).

* Language: expressions whose values the Lua 5.1 reference manual fixes
  (precedence and associativity, %, coercion, and/or, scoping, closures,
  loops, multiple results), each evaluated by a script's Populate.
* The Information / Populate contract of VerifyAndObtainData (:579-605) and
  SetTableFromParamVector (:607-641).
* The scripts under tests/golden/scripted/ against what fixes their rows
  independently: the manual-symmetry ("DOL") rows the reference's ring
  fixtures hold for the same ring, and closed forms of the lattice and helix.
* When the reference tree is present, its own symmetry scripts give the same
  rows as those scripts (skipped elsewhere).
"""
import glob
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SCRIPTED = os.path.join(GOLDEN, "scripted")
REF = "/root/reference"
HEAD = 'Information = { Name = "t", Type = "Symmetry", NLP = 1 }\n'


def rows(script, params=((0,),)):
    import dplus_mi355x as D
    return D.lua_symmetry_rows(script, [list(p) for p in params])


def value(expr, prelude=""):
    """The number a Lua expression evaluates to (row 1, column 1)."""
    return rows(HEAD + prelude + f"\nfunction Populate(p, n) return {{ {{ {expr} }} }} end")[0, 0]


def truth(expr, prelude=""):
    return value(f"(({expr}) and 1 or 0)", prelude) == 1


def mine(name):
    return open(os.path.join(SCRIPTED, name)).read()


# ------------------------------------------------------------------ language
@pytest.mark.parametrize("expr,expected", [
    ("2^3^2", 512.0),           # ^ is right associative
    ("-2^2", -4.0),             # unary minus binds looser than ^
    ("1 + 2 * 3 - 4 / 2", 5.0),
    ("(1 + 2) * 3", 9.0),
    ("7 % 3", 1.0),
    ("-7 % 3", 2.0),            # a % b == a - floor(a / b) * b
    ("7 % -3", -2.0),
    ("5.5 % 2", 1.5),
    ('"10" + 1', 11.0),         # strings coerce to numbers in arithmetic
    ('"0x10" * 1', 16.0),
    ("#'abc'", 3.0),
    ("#{1, 2, 3, nil}", 3.0),
    ("1 < 2 and 10 or 20", 10.0),
    ("nil or 5", 5.0),
    ("false and 1 or 2", 2.0),
    ("math.floor(-3.5)", -4.0),
    ("math.ceil(-3.5)", -3.0),
    ("math.fmod(-7, 3)", -1.0),
    ("math.max(3, 9, 2) + math.min(4, -1)", 8.0),
    ("math.pow(2, 10)", 1024.0),
    ("math.deg(math.pi)", 180.0),
    ("math.atan2(1, -1)", 3 * math.pi / 4),
    ("math.abs(-2.5)", 2.5),
    ("select and 1 or 0", 0.0),  # not in this subset: a nil global
    ("1e3 + .5 + 0x1F", 1031.5),
    ("3 == 3.0 and 1 or 0", 1.0),
])
def test_expression_values(expr, expected):
    assert value(expr) == pytest.approx(expected, rel=1e-15, abs=0)


@pytest.mark.parametrize("expr", [
    '"a" .. 1 .. 2 == "a12"',
    '1.5 .. "" == "1.5"',
    '10 .. "" == "10"',          # integral numbers print without a fraction
    '"abc" < "abd"',
    '2 ~= "2"',                  # no coercion in comparisons for equality
    'not nil',
    'type(nil) == "nil" and type({}) == "table" and type(print) == "function"',
    'tostring(true) == "true"',
    'tonumber("3.25") == 3.25 and tonumber("x") == nil',
    'string.format("%5.2f|%d|%s|%g", 3.14159, 42.9, "ok", 0.5) == " 3.14|42|ok|0.5"',
    'os == nil and io == nil',   # no system access, as in the reference
    '#"" == 0',
])
def test_boolean_expressions(expr):
    assert truth(expr)


def test_statements_and_scoping():
    prelude = """
    local function counter()
      local n = 0
      return function() n = n + 1; return n end
    end
    function fib(n) if n < 2 then return n end return fib(n - 1) + fib(n - 2) end
    function two() return 3, 4 end
    function loops()
      local s = 0
      for i = 1, 2, 0.25 do s = s + 1 end            -- 5 passes
      for i = 10, 1, -3 do s = s + 100 end           -- 10, 7, 4, 1
      for i = 1, 3 do local j = i; i = 99; s = s + j * 1000 end  -- the loop variable is a copy
      local k = 0
      while true do k = k + 1; if k == 7 then break end end
      repeat local stop = k > 9; k = k + 1 until stop  -- until sees the body's locals
      return s + k * 1e6
    end
    obj = { base = 5 }
    function obj:add(x) return self.base + x end
    function swap() local a, b = 1, 2; a, b = b, a; return a * 10 + b end
    function shadow() local x = 1; do local x = 2 end; return x end
    function multi() local a, b, c = two(); local d = (two()); local t = { two(), two() }; return a + b + (c or 100) + d + #t end
    function tables()
      local t = {}
      t[1.0] = "a"; t[2] = "b"; t.x = 7; t["y"] = 8
      table.insert(t, "c"); table.insert(t, 1, "z")
      local keys = 0
      for k, v in pairs(t) do keys = keys + 1 end
      local sum = 0
      for i, v in ipairs({ 4, 5, 6 }) do sum = sum + i * v end
      return (t[1] == "z" and t[4] == "c" and #t == 4 and t.x + t.y == 15 and keys == 6) and sum or -1
    end
    """
    c = "local c = counter(); c(); c(); return { { c(), fib(20), loops(), obj:add(3), swap(), shadow(), multi(), tables() } }"
    r = rows(HEAD + prelude + f"\nfunction Populate(p, n) {c} end")[0]
    # 6 columns are read; the last two (multi, tables) through a second call
    assert list(r) == [3.0, 6765.0, 5 + 400 + 6000 + 11 * 1e6, 8.0, 21.0, 1.0]
    c2 = "return { { multi(), tables() } }"
    r2 = rows(HEAD + prelude + f"\nfunction Populate(p, n) {c2} end")[0]
    assert list(r2[:2]) == [3 + 4 + 100 + 3 + 3, 1 * 4 + 2 * 5 + 3 * 6]


# ------------------------------------------------------------------ the script contract
def test_parameter_table_and_rows():
    """p[i][j] is layer i's parameter j (1-based) for j up to Information.NLP,
    missing entries 0; nlayers is the layer count; row entries that are not
    numbers read as 0 (getlocrot, Symmetry.cpp:437-461)."""
    s = ('Information = { Type = "Symmetry", NLP = 3 }\n'
         'function Populate(p, n) return { { n, #p, #p[1], p[1][1], p[2][2], p[2][3] }, '
         '{ "x", nil, true, 4 }, 7, {} } end')
    r = rows(s, [[1.5], [2, 3.25]])
    assert r.tolist() == [[2, 2, 3, 1.5, 3.25, 0], [0, 0, 0, 4, 0, 0], [0] * 6, [0] * 6]


@pytest.mark.parametrize("script,msg", [
    ('function Populate(p, n) return {} end', "Information"),
    ('Information = { Type = "Geometry", NLP = 1 }\nfunction Populate(p, n) return {} end', "Symmetry"),
    ('Information = { Type = "Symmetry", NLP = 0 }\nfunction Populate(p, n) return {} end', "NLP"),
    ('Information = { Type = "Symmetry", NLP = 1 }', "Populate"),
    (HEAD + 'function Populate(p, n) return 5 end', "table"),
    (HEAD + 'function Populate(p, n) error("Parameter matrix must be 6x1, it is " .. n .. "x" .. 1) end',
     "Parameter matrix must be 6x1, it is 1x1"),
    (HEAD + 'function Populate(p, n) return undefined_name + 1 end', "arithmetic"),
    (HEAD + 'function Populate(p, n) return nofunc() end', "attempt to call"),
    (HEAD + 'function Populate(p, n) return {{ 1 }} ', "'end' expected"),
    (HEAD + 'function Populate(p, n) local t = nil; return t.x end', "attempt to index"),
    (HEAD + 'function Populate(p, n) return { ... } end', "varargs"),
])
def test_script_errors(script, msg):
    import dplus_mi355x as D
    with pytest.raises(D.DplusError) as e:
        rows(script)
    assert e.value.error_code == 21 and msg in str(e.value)


def test_runaway_script_is_stopped():
    import time
    import dplus_mi355x as D
    t = time.time()
    with pytest.raises(D.DplusError, match="statements"):
        rows(HEAD + "function Populate(p, n) while true do end end")
    assert time.time() - t < 30


# ------------------------------------------------------------------ the test scripts
def euler_xyz(a, b, c):
    """Rx(a) Ry(b) Rz(c) (EulerD, Backend/Backend/Amplitude.h:557-573)."""
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rx = np.array([[1, 0, 0], [0, ca, -sa], [0, sa, ca]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rz = np.array([[cc, -sc, 0], [sc, cc, 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


def test_cartesian_lattice_closed_form():
    P = [[3], [4], [2], [4], [3], [8.13]]
    r = rows(mine("cartesian_lattice.lua"), P)
    i, j, k = np.meshgrid(np.arange(3), np.arange(2), np.arange(3), indexing="ij")
    expect = np.stack([4 * i.ravel(), 4 * j.ravel(), 8.13 * k.ravel()], axis=1)
    assert r.shape == (18, 6)
    np.testing.assert_allclose(r[:, :3], expect, rtol=0, atol=1e-12)
    assert not r[:, 3:].any()


@pytest.mark.parametrize("kind", ["Impl", "Expl"])
def test_helical_ring_reproduces_the_reference_dol_rows(kind):
    """The ring fixtures describe one 13-unit ring twice: by the helix script
    with layers (18, 0, 13, 13, 1, 1, 0, 0, 0), and by a manual symmetry whose
    13 rows (6 significant digits) the reference's state holds."""
    scr = json.load(open(os.path.join(SCRIPTED, f"{kind}_Ring_Solution_with_Solvation.json")))
    dol = json.load(open(os.path.join(SCRIPTED, f"{kind}_Ring_DOL_Solution_with_Solvation.json")))
    node = scr["state"]["Domain"]["Populations"][0]["Models"][0]
    r = rows(mine(scr["script"]), node["Parameters"])
    ref = np.array(dol["state"]["Domain"]["Populations"][0]["Models"][0]["Parameters"], dtype=float)
    assert r.shape == ref.shape == (13, 6)
    np.testing.assert_allclose(r, ref, rtol=0, atol=5e-4)


def test_helical_ring_twist_orientation():
    """With an in-ring twist every unit's X-Y-Z Euler angles give the rotation
    Rx(u twa) Rz(90 deg - theta) Ry(u twb); positions lie on the helix."""
    P = [[12], [8], [13], [10], [4], [3], [400], [-20], [15]]
    r = rows(mine("helical_ring.lua"), P)
    assert r.shape == (40, 6)
    rad = np.radians
    step = 8 * 2 / 3
    shift = 2 * np.pi * step / (400 * 13)
    k = 0
    for h in range(4):
        th0 = math.fmod(h * (2 / 3) * shift * 13, 2 * np.pi)
        for u in range(10):
            th = th0 + u * (2 * np.pi / 13 + shift)
            np.testing.assert_allclose(r[k, :3], [12 * np.sin(th), 12 * np.cos(th), h * step + u / 13 * 8],
                                       rtol=0, atol=1e-12)
            Rx = euler_xyz(rad(u * -20), 0, 0)
            Rz = euler_xyz(0, 0, rad(90) - th)
            Ry = euler_xyz(0, rad(u * 15), 0)
            np.testing.assert_allclose(euler_xyz(*rad(r[k, 3:])), Rx @ Rz @ Ry, rtol=0, atol=1e-12)
            k += 1


def test_left_hand_helix_closed_form():
    P = [[11.94], [12.175], [14], [14], [2], [48], [3]]
    r = rows(mine("left_hand_helix.lua"), P)
    h, u = np.meshgrid(np.arange(48), np.arange(2, 14), indexing="ij")
    th = u.ravel() * 2 * np.pi / 14
    xyz = np.stack([11.94 * np.sin(th), 11.94 * np.cos(th), h.ravel() * 2 * 12.175 / 3 + u.ravel() * 12.175 / 14], 1)
    xyz -= xyz.mean(axis=0)
    assert r.shape == (48 * 12, 6)
    np.testing.assert_allclose(r[:, :3], xyz, rtol=0, atol=1e-11)
    np.testing.assert_allclose(r[:, 5], 90 - np.degrees(th), rtol=0, atol=1e-11)
    assert not r[:, 3:5].any()


# ------------------------------------------------------------------ the reference's own scripts
REF_STATES = sorted(glob.glob(os.path.join(REF, "PythonInterface/tests/reviewer_tests/files_for_tests/generate/**/*.state"),
                              recursive=True)) + \
    sorted(glob.glob(os.path.join(REF, "Example Files/**/*.state"), recursive=True))


def _scripted_nodes():
    out, seen = [], set()
    for st in REF_STATES:
        try:
            d = json.load(open(st))
        except (ValueError, UnicodeDecodeError):
            continue

        def walk(m):
            if m.get("Type") == "Scripted Symmetry":
                path = os.path.join(os.path.dirname(st), os.path.basename(m["Filename"].replace("\\", "/")))
                key = (os.path.basename(path), json.dumps(m["Parameters"]))
                if os.path.exists(path) and key not in seen:
                    seen.add(key)
                    out.append((path, m["Parameters"]))
            for c in m.get("Children", []):
                walk(c)
        for p in d.get("Domain", {}).get("Populations", []):
            for m in p.get("Models", []):
                walk(m)
    return out


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference tree is not present")
def test_reference_scripts_match():
    """Every scripted symmetry in the reference's states and examples runs, and
    those with a script under tests/golden/scripted/ give its rows; the 1-D
    rod and 2-D slab lattices follow their closed forms."""
    SCRIPT_FOR = {"Space filling symmetry - Cartesian.lua": "cartesian_lattice.lua",
                  "Super Helical Left Hand MT Helix with in-ring twist.lua": "helical_ring.lua",
                  "Left_Hand_Helix.lua": "left_hand_helix.lua"}
    nodes = _scripted_nodes()
    assert len(nodes) >= 10
    compared = 0
    for path, params in nodes:
        r = rows(open(path).read(), params)
        name = os.path.basename(path)
        if name in SCRIPT_FOR:
            np.testing.assert_allclose(r, rows(mine(SCRIPT_FOR[name]), params), rtol=0, atol=1e-9, err_msg=name)
            compared += 1
        elif name == "10_rod.lua":
            d, n = params[0]
            z = d * np.arange(n) - (n / 2 - 0.5) * d
            np.testing.assert_allclose(r[:, 2], z, rtol=0, atol=1e-12)
            compared += 1
        elif name in ("10_10_Slab.lua", "Grid_in_XY.lua"):
            (dx, nx), (dy, ny) = params
            m, n = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
            np.testing.assert_allclose(r[:, 0], (dx * m - (nx / 2 - 0.5) * dx).ravel(), rtol=0, atol=1e-12)
            np.testing.assert_allclose(r[:, 1], (dy * n - (ny / 2 - 0.5) * dy).ravel(), rtol=0, atol=1e-12)
            compared += 1
        elif name == "MT_Helix.lua":  # turns x (units placed per turn - first unit placed) rows
            v = [row[0] for row in params]
            assert len(r) == int(v[5]) * int(v[3] - v[4]), name
        else:
            assert len(r) > 0, name
    assert compared >= 7
