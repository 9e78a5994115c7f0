reset
set style data lines
set nokey

set xlabel 'x (m)'
set title 'Density_IG_air.xml'
plot "cut1D_CPU0_TIME0.out" u 1:2,\
 "cut1D_CPU0_TIME1.out" u 1:2,\
 "cut1D_CPU0_TIME2.out" u 1:2
pause(-1)

set title 'Pressure_IG_air.xml'
plot "cut1D_CPU0_TIME0.out" u 1:3,\
 "cut1D_CPU0_TIME1.out" u 1:3,\
 "cut1D_CPU0_TIME2.out" u 1:3
pause(-1)

set title 'Temperature_IG_air.xml'
plot "cut1D_CPU0_TIME0.out" u 1:4,\
 "cut1D_CPU0_TIME1.out" u 1:4,\
 "cut1D_CPU0_TIME2.out" u 1:4
pause(-1)

set title 'Velocity_IG_air.xml'
plot "cut1D_CPU0_TIME0.out" u 1:5,\
 "cut1D_CPU0_TIME1.out" u 1:5,\
 "cut1D_CPU0_TIME2.out" u 1:5
pause(-1)

set title 'Niveau AMR'
plot "cut1D_CPU0_TIME0.out" u 1:6,\
 "cut1D_CPU0_TIME1.out" u 1:6,\
 "cut1D_CPU0_TIME2.out" u 1:6
pause(-1)

set title 'Xi'
plot "cut1D_CPU0_TIME0.out" u 1:7,\
 "cut1D_CPU0_TIME1.out" u 1:7,\
 "cut1D_CPU0_TIME2.out" u 1:7
pause(-1)

