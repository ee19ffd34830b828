# par lines (name, radius, charge)
par A 115.238000 0.0
par B 85.229550 0.0
par C 66.488380 0.0
par D 49.157060 0.0
par E 45.433690 0.0
par F 43.058490 0.0
par G 42.479310 0.0
par H 39.162460 0.0
par I 35.943260 0.0
par J 36.761520 0.0
par K 31.712750 0.0
par L 25.749940 0.0
par M 24.011790 0.0
par N 21.424980 0.0
par O 27.771590 0.0
