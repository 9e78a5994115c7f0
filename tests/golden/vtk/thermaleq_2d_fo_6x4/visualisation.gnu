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
splot "result_CPU0_TIME0.out" u 1:2:3,\
 "result_CPU0_TIME1.out" u 1:2:3
pause(-1)

set title 'Density_IG_air.xml'
splot "result_CPU0_TIME0.out" u 1:2:4,\
 "result_CPU0_TIME1.out" u 1:2:4
pause(-1)

set title 'Alpha_SG_water.xml'
splot "result_CPU0_TIME0.out" u 1:2:5,\
 "result_CPU0_TIME1.out" u 1:2:5
pause(-1)

set title 'Density_SG_water.xml'
splot "result_CPU0_TIME0.out" u 1:2:6,\
 "result_CPU0_TIME1.out" u 1:2:6
pause(-1)

set title 'Density_Mixture'
splot "result_CPU0_TIME0.out" u 1:2:7,\
 "result_CPU0_TIME1.out" u 1:2:7
pause(-1)

set title 'Pressure_Mixture'
splot "result_CPU0_TIME0.out" u 1:2:8,\
 "result_CPU0_TIME1.out" u 1:2:8
pause(-1)

set title 'Temperature_Mixture'
splot "result_CPU0_TIME0.out" u 1:2:9,\
 "result_CPU0_TIME1.out" u 1:2:9
pause(-1)

set title 'Velocity_Mixture_norm'
splot "result_CPU0_TIME0.out" u 1:2:10,\
 "result_CPU0_TIME1.out" u 1:2:10
pause(-1)

set title 'Niveau AMR'
splot "result_CPU0_TIME0.out" u 1:2:11,\
 "result_CPU0_TIME1.out" u 1:2:11
pause(-1)

set title 'Xi'
splot "result_CPU0_TIME0.out" u 1:2:12,\
 "result_CPU0_TIME1.out" u 1:2:12
pause(-1)

