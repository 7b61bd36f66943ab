!! Readers of the wavefunction and basis inputs, with the behaviour of the
!! reference's src/mod_read_files.f90: filewfx (:12-287), filefchk
!! (:432-691) and filebas (:695-726).  Primitives are expanded to Cartesian
!! components in the reference's component order.
module read_files
   use geninfo
   use wfxinfo
   use locatemod
   implicit none

   ! Cartesian components per angular momentum, in the order the reference
   ! assigns them (Ptyp 1..35)
   integer, parameter :: ncomp(0:4) = (/1, 3, 6, 10, 15/)
   integer, parameter :: comp_first(0:4) = (/1, 2, 5, 11, 21/)
   integer, parameter :: comp_tmn(3, 35) = reshape((/ &
      0,0,0, &
      1,0,0, 0,1,0, 0,0,1, &
      2,0,0, 0,2,0, 0,0,2, 1,1,0, 1,0,1, 0,1,1, &
      3,0,0, 0,3,0, 0,0,3, 2,1,0, 2,0,1, 0,2,1, 1,2,0, 1,0,2, 0,1,2, 1,1,1, &
      4,0,0, 0,4,0, 0,0,4, 3,1,0, 3,0,1, 1,3,0, 0,3,1, 1,0,3, 0,1,3, 2,2,0, &
      2,0,2, 0,2,2, 2,1,1, 1,2,1, 1,1,2 /), (/3, 35/))

contains

   !****************************************************************************
   !! .wfx reader (src/mod_read_files.f90:12-287)
   subroutine filewfx(wfxfilename)
      character(len=*), intent(in) :: wfxfilename
      character(len=80) :: line
      integer :: i, j, kk, mon
      double precision :: maxim, minim
      double precision, parameter :: trsh = 1.d-16

      corr = .false.; uhf = .false.; rhf = .false.; udens = .false.; rdens = .false.
      opsh = .false.; clsh = .false.
      open(unit=1, file=wfxfilename, status='OLD')
      call locate(1, "Number of Nuclei");        read(1,*) natoms
      call locate(1, "Number of Occupied Molecular Orbitals"); read(1,*) noccmo
      allocate(Occ(noccmo))
      call locate(1, "Molecular Orbital Occupation Numbers")
      do i = 1, noccmo
         read(1,*) Occ(i)
      end do
      do i = 1, noccmo
         if (Occ(i) /= 0.d0 .and. Occ(i) /= 1.d0 .and. Occ(i) /= 2.d0) then
            write(*,*) "Correlated wavefunction"
            corr = .true.
            exit
         end if
      end do
      maxim = maxval(Occ); minim = minval(Occ)
      if (corr) then
         if (maxim >= 2.d0) then
            if ((maxim - 2.d0 > trsh) .or. (minim < 0.d0)) then
               write(*,*) "RELAXED DENSITY (closed shell)"; rdens = .true.
            else
               write(*,*) "UNRELAXED DENSITY (closed shell)"; udens = .true.
            end if
         else
            if ((maxim - 1.d0 > trsh) .or. (minim < 0.d0)) then
               write(*,*) "RELAXED DENSITY (open shell)"; rdens = .true.
            else
               write(*,*) "UNRELAXED DENSITY (open shell)"; udens = .true.
            end if
         end if
      else
         minim = minval(Occ, mask=Occ > 0.d0)
         if (minim == 1.d0) then
            opsh = .true.
            if (maxim == 2.d0) then
               rhf = .true.; write(*,*) "Open shell RHF"
            else
               uhf = .true.; write(*,*) "UHF"
            end if
         else if (minim == 2.d0) then
            rhf = .true.; clsh = .true.; write(*,*) "Closed shell RHF"
         else
            write(*,*) "Error, correlated wf"
            stop
         end if
      end if
      call locate(1, "Net Charge");                read(1,*) netch
      call locate(1, "Number of Electrons");       read(1,*) nelec
      call locate(1, "Number of Alpha Electrons"); read(1,*) nalfae
      call locate(1, "Number of Beta Electrons");  read(1,*) nbetae
      call locate(1, "Number of Primitives");      read(1,*) nprim
      allocate(cartes(natoms,3))
      call locate(1, "Nuclear Cartesian Coordinates")
      read(1,*) ((cartes(i,j), j=1,3), i=1,natoms)
      allocate(Ra(nprim), Ptyp(nprim), Alpha(nprim))
      call locate(1, "Primitive Centers");   read(1,*) (Ra(i), i=1,nprim)
      call locate(1, "Primitive Types");     read(1,*) (Ptyp(i), i=1,nprim)
      call locate(1, "Primitive Exponents"); read(1,*) (Alpha(i), i=1,nprim)
      allocate(an(natoms), chrg(natoms))
      call locate(1, "Atomic Numbers")
      do i = 1, natoms
         read(1,*) an(i)
      end do
      call locate(1, "Nuclear Charges")
      do i = 1, natoms
         read(1,*) chrg(i)
      end do
      allocate(T(noccmo, nprim))
      call locate(1, "Molecular Orbital Primitive Coefficients")
      do kk = 1, noccmo
         read(1,'(a80)') line
         read(1,*) mon
         read(1,'(a80)') line
         read(1,*) (T(mon,i), i=1,nprim)
      end do
      allocate(TMN(nprim,3))
      do i = 1, nprim
         if (Ptyp(i) < 1 .or. Ptyp(i) > 20) then
            write(*,*) "G orbitals are not implemented"
            stop
         end if
         TMN(i,:) = comp_tmn(:, Ptyp(i))
      end do
      allocate(Xn(nprim), Yn(nprim), Zn(nprim))
      do i = 1, nprim
         Xn(i) = cartes(Ra(i),1); Yn(i) = cartes(Ra(i),2); Zn(i) = cartes(Ra(i),3)
      end do
      close(1)
   end subroutine filewfx

   !****************************************************************************
   !! .fchk reader (src/mod_read_files.f90:432-691).  Shell types 0, 1, +-2,
   !! +-3, +-4 (pure shells are expanded as Cartesian, as the reference does)
   !! and -1 (SP: all s primitives, then all px, py, pz).
   subroutine filefchk(fchkfilename)
      character(len=*), intent(in) :: fchkfilename
      character(len=500) :: line
      character(len=40) :: typ, method, basis
      integer :: i, k, l, c, L0, iprim, ncshl, npshl, st
      integer, allocatable :: shell_types(:), prim_per_shell(:), shell2atom(:)
      double precision, allocatable :: prim_exp(:)

      corr = .false.; uhf = .false.; rhf = .false.; udens = .false.; rdens = .false.
      opsh = .false.; clsh = .false.
      open(unit=1, file=fchkfilename, status='OLD')
      call getline_with(1, "SP", line)
      read(line,*) typ, method, basis
      call getline_with(1, "Number of alpha electrons", line)
      read(line,*) typ, typ, typ, typ, typ, nalfae
      call getline_with(1, "Number of beta electrons", line)
      read(line,*) typ, typ, typ, typ, typ, nbetae
      nelec = nalfae + nbetae
      call getline_with(1, "Total Energy", line)
      read(line,*) typ, typ, typ, toteng
      write(*,*) "Total Energy:", toteng
      write(*,*) "Calculation type:", trim(method), trim(basis)
      if (nalfae == nbetae) then
         clsh = .true.; write(*,*) "Closed shell"
      else
         opsh = .true.; write(*,*) "Open shell"
      end if
      select case (trim(method))
      case ("RHF")
         rhf = .true.
      case ("HF")
         if (opsh) uhf = .true.
         if (clsh) rhf = .true.
      case ("UHF")
         write(*,*) "Open shell UHF"; uhf = .true.
      case default
         corr = .true.; write(*,*) "Correlated wavefunction"
      end select
      call getline_with(1, "Number of atoms", line)
      read(line,*) typ, typ, typ, typ, natoms
      allocate(an(natoms), chrg(natoms), cartes(natoms,3))
      call locate(1, "Atomic numbers");          read(1,*) (an(i), i=1,natoms)
      chrg(:) = dble(an(:))
      call locate(1, "Current cartesian coord"); read(1,*) ((cartes(i,k), k=1,3), i=1,natoms)
      call getline_with(1, "Number of contracted shells", line)
      read(line,*) typ, typ, typ, typ, typ, ncshl
      call getline_with(1, "Number of primitive shells", line)
      read(line,*) typ, typ, typ, typ, typ, npshl
      allocate(shell_types(ncshl), prim_per_shell(ncshl), shell2atom(ncshl), prim_exp(npshl))
      call locate(1, "Shell types");                    read(1,*) (shell_types(i), i=1,ncshl)
      call locate(1, "Number of primitives per shell"); read(1,*) (prim_per_shell(i), i=1,ncshl)
      nprim = 0
      do i = 1, ncshl
         st = shell_types(i)
         if (st == -1) then
            nprim = nprim + 4*prim_per_shell(i)
         else if (abs(st) <= 4) then
            nprim = nprim + ncomp(abs(st))*prim_per_shell(i)
         else
            write(*,*) "Shell type not implemented:", st
            stop
         end if
      end do
      call locate(1, "Shell to atom map");   read(1,*) (shell2atom(i), i=1,ncshl)
      call locate(1, "Primitive exponents"); read(1,*) (prim_exp(i), i=1,npshl)
      allocate(Ra(nprim), Alpha(nprim), Ptyp(nprim), TMN(nprim,3))
      iprim = 0
      l = 0
      do i = 1, ncshl
         st = shell_types(i)
         if (st == -1) then
            do L0 = 0, 3        ! s block, then px, py, pz blocks
               do k = 1, prim_per_shell(i)
                  iprim = iprim + 1
                  Ra(iprim) = shell2atom(i)
                  Alpha(iprim) = prim_exp(l + k)
                  Ptyp(iprim) = L0 + 1
                  TMN(iprim,:) = comp_tmn(:, L0 + 1)
               end do
            end do
            l = l + prim_per_shell(i)
         else
            do k = 1, prim_per_shell(i)
               l = l + 1
               do c = 0, ncomp(abs(st)) - 1
                  iprim = iprim + 1
                  Ra(iprim) = shell2atom(i)
                  Alpha(iprim) = prim_exp(l)
                  Ptyp(iprim) = comp_first(abs(st)) + c
                  TMN(iprim,:) = comp_tmn(:, comp_first(abs(st)) + c)
               end do
            end do
         end if
      end do
      close(1)
   end subroutine filefchk

   !****************************************************************************
   !! .bas reader (src/mod_read_files.f90:695-726): after the first line with
   !! '#', nprim lines "k x y z alpha t m n" override the primitive centres,
   !! exponents and t,m,n; Ra is the atom whose coordinates match within 1e-5.
   subroutine filebas(basfilename)
      character(len=*), intent(in) :: basfilename
      integer :: i, j, k
      double precision :: b(3)
      open(unit=1, file=basfilename, status='OLD')
      if (allocated(Xn)) deallocate(Xn, Yn, Zn)
      allocate(Xn(nprim), Yn(nprim), Zn(nprim))
      call locate(1, "#")
      do i = 1, nprim
         read(1,*) k, b(1), b(2), b(3), Alpha(i), TMN(i,1), TMN(i,2), TMN(i,3)
         Xn(i) = b(1); Yn(i) = b(2); Zn(i) = b(3)
         do j = 1, natoms
            if (abs(b(1) - cartes(j,1)) < 0.00001 .and. abs(b(2) - cartes(j,2)) < 0.00001 &
                .and. abs(b(3) - cartes(j,3)) < 0.00001) Ra(i) = j
         end do
      end do
      close(1)
   end subroutine filebas

end module read_files
