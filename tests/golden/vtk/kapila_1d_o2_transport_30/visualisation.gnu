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

set title 'Temperature_IG_air.xml'
plot "result_CPU0_TIME0.out" u 1:4,\
 "result_CPU0_TIME1.out" u 1:4
pause(-1)

set title 'Mass fraction_IG_air.xml'
plot "result_CPU0_TIME0.out" u 1:5,\
 "result_CPU0_TIME1.out" u 1:5
pause(-1)

set title 'Alpha_SG_water.xml'
plot "result_CPU0_TIME0.out" u 1:6,\
 "result_CPU0_TIME1.out" u 1:6
pause(-1)

set title 'Density_SG_water.xml'
plot "result_CPU0_TIME0.out" u 1:7,\
 "result_CPU0_TIME1.out" u 1:7
pause(-1)

set title 'Temperature_SG_water.xml'
plot "result_CPU0_TIME0.out" u 1:8,\
 "result_CPU0_TIME1.out" u 1:8
pause(-1)

set title 'Mass fraction_SG_water.xml'
plot "result_CPU0_TIME0.out" u 1:9,\
 "result_CPU0_TIME1.out" u 1:9
pause(-1)

set title 'Density_Mixture'
plot "result_CPU0_TIME0.out" u 1:10,\
 "result_CPU0_TIME1.out" u 1:10
pause(-1)

set title 'Pressure_Mixture'
plot "result_CPU0_TIME0.out" u 1:11,\
 "result_CPU0_TIME1.out" u 1:11
pause(-1)

set title 'Velocity_Mixture'
plot "result_CPU0_TIME0.out" u 1:12,\
 "result_CPU0_TIME1.out" u 1:12
pause(-1)

set title 'Transport1'
plot "result_CPU0_TIME0.out" u 1:13,\
 "result_CPU0_TIME1.out" u 1:13
pause(-1)

set title 'Transport2'
plot "result_CPU0_TIME0.out" u 1:14,\
 "result_CPU0_TIME1.out" u 1:14
pause(-1)

set title 'Niveau AMR'
plot "result_CPU0_TIME0.out" u 1:15,\
 "result_CPU0_TIME1.out" u 1:15
pause(-1)

set title 'Xi'
plot "result_CPU0_TIME0.out" u 1:16,\
 "result_CPU0_TIME1.out" u 1:16
pause(-1)

