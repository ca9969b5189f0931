Basic divider and sources
* a comment line
R1 in mid 1k
R2 mid 0 2K ; inline comment
Vdd in 0 DC 5
I1 0 mid 1m
C1 mid gnd 10p
.op
.end
R3 after end 1
