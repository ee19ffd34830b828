# par lines (name, radius, charge)
