-- A rectangular lattice of copies: layers are (count x, spacing x, count y,
-- spacing y, count z, spacing z); copies are listed with z varying fastest.
Information = { Name = "Cartesian lattice", Type = "Symmetry", NLP = 1, MinLayers = 6, MaxLayers = 6 }

local function axis(p, k)
  return p[2 * k - 1][1], p[2 * k][1]
end

function Populate(p, nlayers)
  if nlayers ~= 6 then
    error("a Cartesian lattice takes 6 layers, got " .. nlayers)
  end
  local nx, dx = axis(p, 1)
  local ny, dy = axis(p, 2)
  local nz, dz = axis(p, 3)
  local rows = {}
  local i = 0
  while i < nx do
    for j = 0, ny - 1 do
      for k = 0, nz - 1 do
        table.insert(rows, { i * dx, j * dy, k * dz, 0, 0, 0 })
      end
    end
    i = i + 1
  end
  return rows
end
