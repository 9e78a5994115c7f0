reset
set style data lines
set nokey

set xlabel 'x (m)'
set title 'Alpha_IG_air.xml'
plot "result_CPU0_TIME0.out" u 1:2,\
 "result_CPU0_TIME1.out" u 1:2
pause(-1)

set title 'Density_IG_air.xml'
plot "result_CPU0_TIME0.out" u 1:3,\
 "result_CPU0_TIME1.out" u 1:3
pause(-1)

set title 'Alpha_SG_water.xml'
plot "result_CPU0_TIME0.out" u 1:4,\
 "result_CPU0_TIME1.out" u 1:4
pause(-1)

set title 'Density_SG_water.xml'
plot "result_CPU0_TIME0.out" u 1:5,\
 "result_CPU0_TIME1.out" u 1:5
pause(-1)

set title 'Density_Mixture'
plot "result_CPU0_TIME0.out" u 1:6,\
 "result_CPU0_TIME1.out" u 1:6
pause(-1)

set title 'Pressure_Mixture'
plot "result_CPU0_TIME0.out" u 1:7,\
 "result_CPU0_TIME1.out" u 1:7
pause(-1)

set title 'Temperature_Mixture'
plot "result_CPU0_TIME0.out" u 1:8,\
 "result_CPU0_TIME1.out" u 1:8
pause(-1)

set title 'Velocity_Mixture_norm'
plot "result_CPU0_TIME0.out" u 1:9,\
 "result_CPU0_TIME1.out" u 1:9
pause(-1)

set title 'Niveau AMR'
plot "result_CPU0_TIME0.out" u 1:10,\
 "result_CPU0_TIME1.out" u 1:10
pause(-1)

set title 'Xi'
plot "result_CPU0_TIME0.out" u 1:11,\
 "result_CPU0_TIME1.out" u 1:11
pause(-1)

