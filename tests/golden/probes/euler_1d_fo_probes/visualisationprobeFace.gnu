reset
set style data lines
set nokey

set xlabel 't (s)'
set title 'Density_IG_air.xml'
plot "probeFace.out" u 1:2
pause(-1)

set title 'Pressure_IG_air.xml'
plot "probeFace.out" u 1:3
pause(-1)

set title 'Temperature_IG_air.xml'
plot "probeFace.out" u 1:4
pause(-1)

set title 'Velocity_IG_air.xml'
plot "probeFace.out" u 1:5
pause(-1)

set title 'Niveau AMR'
plot "probeFace.out" u 1:6
pause(-1)

set title 'Xi'
plot "probeFace.out" u 1:7
pause(-1)

