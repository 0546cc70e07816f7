-- A left-handed helix of units (radius, pitch, units per turn, units placed per
-- turn, first unit placed, number of turns, helix starts), each unit turned by
-- 90 - theta degrees about z, the whole set centred on its mean position.
Information = { Name = "Left-handed helix", Type = "Symmetry", NLP = 1, MinLayers = 7, MaxLayers = 7 }

function Populate(p, nlayers)
  assert(nlayers == 7, "the helix takes 7 layers")
  local radius, pitch, per_turn, placed, first, turns, starts =
    p[1][1], p[2][1], p[3][1], p[4][1], p[5][1], p[6][1], p[7][1]
  local rise = 2 * pitch / starts
  local rows, mean = {}, { 0, 0, 0 }
  for h = 0, turns - 1 do
    for u = first, placed - 1 do
      local theta = u * 2 * math.pi / per_turn
      local row = { radius * math.sin(theta), radius * math.cos(theta), h * rise + u * pitch / per_turn,
                    0, 0, 90 - math.deg(theta) }
      rows[#rows + 1] = row
      for c = 1, 3 do mean[c] = mean[c] + row[c] end
    end
  end
  for _, row in ipairs(rows) do
    for c = 1, 3 do row[c] = row[c] - mean[c] / #rows end
  end
  return rows
end
