-- Units on a (super-)helix of radius r: `per_turn` units per turn, `in_turn` of
-- them placed per turn, `turns` turns, `starts` helix starts, an optional
-- super-helical pitch that shifts each unit's angle, and an in-ring twist (alpha
-- and beta steps per unit, degrees).  Each unit's orientation is
-- Rx(a) Rz(g) Ry(b) (first beta, then gamma, then alpha), reported as the
-- alpha, beta, gamma of an X-Y-Z Euler rotation.
Information = { Name = "Helical ring with twist", Type = "Symmetry", NLP = 1, MinLayers = 9, MaxLayers = 9 }

local function rx(a) local c, s = math.cos(a), math.sin(a); return { {1, 0, 0}, {0, c, -s}, {0, s, c} } end
local function ry(a) local c, s = math.cos(a), math.sin(a); return { {c, 0, s}, {0, 1, 0}, {-s, 0, c} } end
local function rz(a) local c, s = math.cos(a), math.sin(a); return { {c, -s, 0}, {s, c, 0}, {0, 0, 1} } end

local function mul(A, B)
  local C = {}
  for i = 1, 3 do
    C[i] = {}
    for j = 1, 3 do
      local acc = 0
      for k = 1, 3 do acc = acc + A[i][k] * B[k][j] end
      C[i][j] = acc
    end
  end
  return C
end

-- R = Rx(alpha) Ry(beta) Rz(gamma): R[1][3] = sin(beta), R[2][3] = -cos(beta) sin(alpha),
-- R[3][3] = cos(alpha) cos(beta), R[1][2] = -cos(beta) sin(gamma), R[1][1] = cos(beta) cos(gamma)
local function xyz_angles(R)
  local sb = R[1][3]
  if math.abs(sb) < 1 - 1e-5 then
    local b = math.asin(sb)
    local cb = math.cos(b)
    return math.atan2(-R[2][3] / cb, R[3][3] / cb), b, math.atan2(-R[1][2] / cb, R[1][1] / cb)
  end
  -- gimbal lock: gamma is taken as 0
  local b = sb > 0 and math.pi / 2 or -math.pi / 2
  return math.atan2(R[3][2], R[2][2]), b, 0
end

function Populate(p, nlayers)
  if nlayers ~= 9 then error("the helical ring takes 9 layers, got " .. nlayers) end
  local v = {}
  for l = 1, 9 do v[l] = p[l][1] end
  local r, pitch, per_turn, in_turn, turns, starts, super, twa, twb = unpack(v)
  local step = pitch * 2 / starts
  local dtheta = 2 * math.pi / per_turn
  local shift = 0
  if super > 1e-5 then shift = 2 * math.pi * step / (super * per_turn) end
  local deg = 180 / math.pi
  local rows = {}
  for h = 0, turns - 1 do
    local theta0 = math.fmod(h * (2 / starts) * shift * per_turn, 2 * math.pi)
    for u = 0, in_turn - 1 do
      local theta = theta0 + u * (dtheta + shift)
      local R = mul(rx(u * twa / deg), mul(rz((90 - theta * deg) / deg), ry(u * twb / deg)))
      local a, b, g = xyz_angles(R)
      rows[#rows + 1] = { r * math.sin(theta), r * math.cos(theta), h * step + u / per_turn * pitch,
                          a * deg, b * deg, g * deg }
    end
  end
  return rows
end
