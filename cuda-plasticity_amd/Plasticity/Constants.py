"""Component labels and the Levi-Civita symbol (Plasticity/Constants.py:1-22)."""
x = 'x'
y = 'y'
z = 'z'
xyz = [x, y, z]

perm = {}
for _i in xyz:
    for _j in xyz:
        for _k in xyz:
            perm[_i, _j, _k] = 0.
perm[x, y, z] = perm[y, z, x] = perm[z, x, y] = 1.
perm[x, z, y] = perm[z, y, x] = perm[y, x, z] = -1.
