reset
set style data lines
set nokey

set xlabel 'x (m)'
set surface
set dgrid3d 50,50
set contour base
set cntrparam levels 25
show contour
set title 'Alpha_IG_air.xml'
splot "cut2D_CPU0_TIME0.out" u 1:2:3,\
 "cut2D_CPU0_TIME1.out" u 1:2:3
pause(-1)

set title 'Density_IG_air.xml'
splot "cut2D_CPU0_TIME0.out" u 1:2:4,\
 "cut2D_CPU0_TIME1.out" u 1:2:4
pause(-1)

set title 'Temperature_IG_air.xml'
splot "cut2D_CPU0_TIME0.out" u 1:2:5,\
 "cut2D_CPU0_TIME1.out" u 1:2:5
pause(-1)

set title 'Mass fraction_IG_air.xml'
splot "cut2D_CPU0_TIME0.out" u 1:2:6,\
 "cut2D_CPU0_TIME1.out" u 1:2:6
pause(-1)

set title 'Alpha_SG_water.xml'
splot "cut2D_CPU0_TIME0.out" u 1:2:7,\
 "cut2D_CPU0_TIME1.out" u 1:2:7
pause(-1)

set title 'Density_SG_water.xml'
splot "cut2D_CPU0_TIME0.out" u 1:2:8,\
 "cut2D_CPU0_TIME1.out" u 1:2:8
pause(-1)

set title 'Temperature_SG_water.xml'
splot "cut2D_CPU0_TIME0.out" u 1:2:9,\
 "cut2D_CPU0_TIME1.out" u 1:2:9
pause(-1)

set title 'Mass fraction_SG_water.xml'
splot "cut2D_CPU0_TIME0.out" u 1:2:10,\
 "cut2D_CPU0_TIME1.out" u 1:2:10
pause(-1)

set title 'Density_Mixture'
splot "cut2D_CPU0_TIME0.out" u 1:2:11,\
 "cut2D_CPU0_TIME1.out" u 1:2:11
pause(-1)

set title 'Pressure_Mixture'
splot "cut2D_CPU0_TIME0.out" u 1:2:12,\
 "cut2D_CPU0_TIME1.out" u 1:2:12
pause(-1)

set title 'Velocity_Mixture'
splot "cut2D_CPU0_TIME0.out" u 1:2:13,\
 "cut2D_CPU0_TIME1.out" u 1:2:13
pause(-1)

set title 'Niveau AMR'
splot "cut2D_CPU0_TIME0.out" u 1:2:14,\
 "cut2D_CPU0_TIME1.out" u 1:2:14
pause(-1)

set title 'Xi'
splot "cut2D_CPU0_TIME0.out" u 1:2:15,\
 "cut2D_CPU0_TIME1.out" u 1:2:15
pause(-1)

