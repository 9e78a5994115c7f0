reset
set style data lines
set nokey

set xlabel 'x (m)'
set title 'Alpha_IG_air.xml'
plot "cut1D_CPU0_TIME0.out" u 1:2,\
 "cut1D_CPU0_TIME1.out" u 1:2
pause(-1)

set title 'Density_IG_air.xml'
plot "cut1D_CPU0_TIME0.out" u 1:3,\
 "cut1D_CPU0_TIME1.out" u 1:3
pause(-1)

set title 'Temperature_IG_air.xml'
plot "cut1D_CPU0_TIME0.out" u 1:4,\
 "cut1D_CPU0_TIME1.out" u 1:4
pause(-1)

set title 'Mass fraction_IG_air.xml'
plot "cut1D_CPU0_TIME0.out" u 1:5,\
 "cut1D_CPU0_TIME1.out" u 1:5
pause(-1)

set title 'Alpha_SG_water.xml'
plot "cut1D_CPU0_TIME0.out" u 1:6,\
 "cut1D_CPU0_TIME1.out" u 1:6
pause(-1)

set title 'Density_SG_water.xml'
plot "cut1D_CPU0_TIME0.out" u 1:7,\
 "cut1D_CPU0_TIME1.out" u 1:7
pause(-1)

set title 'Temperature_SG_water.xml'
plot "cut1D_CPU0_TIME0.out" u 1:8,\
 "cut1D_CPU0_TIME1.out" u 1:8
pause(-1)

set title 'Mass fraction_SG_water.xml'
plot "cut1D_CPU0_TIME0.out" u 1:9,\
 "cut1D_CPU0_TIME1.out" u 1:9
pause(-1)

set title 'Density_Mixture'
plot "cut1D_CPU0_TIME0.out" u 1:10,\
 "cut1D_CPU0_TIME1.out" u 1:10
pause(-1)

set title 'Pressure_Mixture'
plot "cut1D_CPU0_TIME0.out" u 1:11,\
 "cut1D_CPU0_TIME1.out" u 1:11
pause(-1)

set title 'Velocity_Mixture'
plot "cut1D_CPU0_TIME0.out" u 1:12,\
 "cut1D_CPU0_TIME1.out" u 1:12
pause(-1)

set title 'Niveau AMR'
plot "cut1D_CPU0_TIME0.out" u 1:13,\
 "cut1D_CPU0_TIME1.out" u 1:13
pause(-1)

set title 'Xi'
plot "cut1D_CPU0_TIME0.out" u 1:14,\
 "cut1D_CPU0_TIME1.out" u 1:14
pause(-1)

