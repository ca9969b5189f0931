Lines the parser treats specially
  * an indented star is not a comment
R1 a b
R2 a 0 1k
V1 a 0
.param vsup=1.8 rload=2k
.option reltol=1e-3
Bfoo a 0 V=1
C1 a 0 1p m=2
.model empty NMOS
.model Nch2 nmos (kp=90u)
.model nch2 NMOS (KP=95u)
M1 a b 0 0 Nch2 W 3u L 1u
M2 b a 0 0 nch2
+ W=5u
R3 b GROUND 1k
R4 b Vss 2k
Iac 0 a AC 1
Vac2 b 0 AC 2 PULSE(0 1 0 1n 1n 5n 0)
.ac dec 10 1 1g
.dc V1 0 1 0.1
.end
