! ISO_C_BINDING interface of libcmaq_ros3.so (include/cmaq_ros3.h) for the
! Fortran CCTM: the binding a gas/ros3_gpu CHEM driver uses in place of the
! per-block CALC_RCONST + RBSOLVER loop of CCTM/src/gas/ros3/rbdriver.F:444-677.
!
! Every derived type below is BIND(C) and matches the C struct of the same
! name field for field (sizes and offsets are checked against the header by
! tests/test_fortran_binding.py).  Arrays go by reference in Fortran order:
! a cell-fastest Y(NCELLS, NUMB_MECH_SPC) is the C layout [species][ncells].
! Mechanism tables are passed as C_LOC addresses of TARGET copies of the
! RXNS_DATA arrays, as tests/fortran/ros3_fdriver.F90 does.
!
! CMAQ_ROS3_WRITE_FAILURE writes to a Fortran unit what RBSOLVER writes to
! LOGDEV before it stops the run (rbsolver.F:532-583), from the library's
! failure record (cmaq_ros3_get_fail_detail); the caller then calls M3EXIT.

module cmaq_ros3_c
   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: CMAQ_ROS3_OK = 0
   integer(c_int), parameter :: CMAQ_ROS3_E_ARG = -1
   integer(c_int), parameter :: CMAQ_ROS3_E_MECH = -2
   integer(c_int), parameter :: CMAQ_ROS3_E_HIP = -3
   integer(c_int), parameter :: CMAQ_ROS3_E_CONVERGENCE = -4
   integer(c_int), parameter :: CMAQ_ROS3_E_STEPS = -5
   integer(c_int), parameter :: CMAQ_ROS3_E_NOMEM = -6
   integer(c_int), parameter :: CMAQ_ROS3_E_TAPE = -7

   ! cmaq_ros3_mech: the RXNS_DATA tables (MECHS/<mech>/RXNS_DATA_MODULE.F90)
   type, bind(c) :: cmaq_ros3_mech_t
      integer(c_int32_t) :: numb_mech_spc, nrxns, mxprd, kunits
      type(c_ptr)        :: irr, sc, nreact, nprdct, ktype, irxbits, iorder, rtdat
      integer(c_int32_t) :: nfalloff
      type(c_ptr)        :: irrfall, rfdat
      integer(c_int32_t) :: nmphot, nphotab
      type(c_ptr)        :: iph
      integer(c_int32_t) :: mhetero, nhetero
      type(c_ptr)        :: ihetero
      integer(c_int32_t) :: nwm, nww, nwo2, nwn2, nwch4, nwh2
      type(c_ptr)        :: nrxwm, nrxww, nrxwo2, nrxwn2, nrxwch4, nrxwh2
      real(c_double)     :: atm_air, atm_o2, atm_n2, atm_ch4, atm_h2
      integer(c_int32_t) :: nspecial_rxn
   end type cmaq_ros3_mech_t

   ! cmaq_ros3_opts
   type, bind(c) :: cmaq_ros3_opts_t
      integer(c_int32_t) :: device = -1        ! HIP device; -1 = current
      integer(c_int32_t) :: max_attempts = 0   ! per cell and call; 0 = 1000000
      integer(c_int32_t) :: waves_per_cu = 0   ! 0 = occupancy limit
      integer(c_int32_t) :: refill_min = 0     ! 0 = autotuned
   end type cmaq_ros3_opts_t

   ! cmaq_ros3_fail
   type, bind(c) :: cmaq_ros3_fail_t
      integer(c_int32_t) :: code, pad
      integer(c_int64_t) :: cell               ! 0-based, -1 if none
      real(c_double)     :: maxerr, dt, tnow
   end type cmaq_ros3_fail_t

   ! cmaq_ros3_fail_detail
   type, bind(c) :: cmaq_ros3_fail_detail_t
      integer(c_int32_t) :: code, numb_mech_spc, nrxns, nphotab
      integer(c_int64_t) :: cell
      real(c_double)     :: maxerr, err, dt, tnow
      real(c_double)     :: temp_k, pres_atm, h2o_ppm, seawater, dens
   end type cmaq_ros3_fail_detail_t

   ! cmaq_ros3_grid: the arrays of one CHEM call (rbdriver.F:20)
   type, bind(c) :: cmaq_ros3_grid_t
      integer(c_int32_t) :: ncols, nrows, nlays, ld_het
      type(c_ptr)        :: cgrid, ta, dens, pres, qv, seawater_zone, seaice, rj, khetero
   end type cmaq_ros3_grid_t

   ! cmaq_gear_tol: GEAR_RTOL, GEAR_MAX_ATOL, GEAR_MIN_ATOL, GEAR_CONC_FLOOR
   ! (REAL(4), RUNTIME_VARS.F:173-176)
   type, bind(c) :: cmaq_gear_tol_t
      real(c_float) :: rtol = 1.0e-3, max_atol = 1.0e-7, min_atol = 1.0e-9, conc_floor = 1.0e-12
   end type cmaq_gear_tol_t

   interface
      ! RBINIT + RBSPARSE (rbinit.F:26, rbsparse.F:26)
      integer(c_int) function cmaq_ros3_init(ctx, mech, rtol, atol, opts) &
            bind(c, name='cmaq_ros3_init')
         import :: c_int, c_ptr, c_double, cmaq_ros3_mech_t, cmaq_ros3_opts_t
         type(c_ptr), intent(out)             :: ctx
         type(cmaq_ros3_mech_t), intent(in)   :: mech
         real(c_double), intent(in)           :: rtol(*), atol(*)
         type(cmaq_ros3_opts_t), intent(in)   :: opts
      end function cmaq_ros3_init

      ! CALC_RCONST + RBSOLVER over all cells (rbdriver.F:444-677), host arrays
      integer(c_int) function cmaq_ros3_run(ctx, ncells, chemstep, temp, pres, h2o, rj, het, &
            seawater, y, nacc, nrej, fail) bind(c, name='cmaq_ros3_run')
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double, cmaq_ros3_fail_t
         type(c_ptr), value                 :: ctx
         integer(c_int64_t), value          :: ncells
         real(c_double), value              :: chemstep            ! minutes
         real(c_double), intent(in)         :: temp(*), pres(*), h2o(*)   ! K, atm, ppm
         real(c_double), intent(in)         :: rj(*)               ! (NCELLS, NPHOTAB) 1/min
         type(c_ptr), value                 :: het, seawater       ! C_LOC(...) or C_NULL_PTR (= 0)
         real(c_double), intent(inout)      :: y(*)                ! (NCELLS, NUMB_MECH_SPC) ppm
         integer(c_int32_t), intent(out)    :: nacc(*), nrej(*)
         type(cmaq_ros3_fail_t), intent(out) :: fail
      end function cmaq_ros3_run

      ! SMVGEAR's tolerances (GRVARS.F:550-707 reads them)
      integer(c_int) function cmaq_gear_set_tolerances(ctx, tol) bind(c, name='cmaq_gear_set_tolerances')
         import :: c_int, c_ptr, cmaq_gear_tol_t
         type(c_ptr), value                 :: ctx
         type(cmaq_gear_tol_t), intent(in)  :: tol
      end function cmaq_gear_set_tolerances

      ! CALC_RCONST + SMVGEAR over all cells (grdriver.F:400-560), host arrays
      integer(c_int) function cmaq_gear_run(ctx, ncells, chemstep, temp, pres, h2o, rj, het, &
            seawater, y, nsteps, nfail, fail) bind(c, name='cmaq_gear_run')
         import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double, cmaq_ros3_fail_t
         type(c_ptr), value                 :: ctx
         integer(c_int64_t), value          :: ncells
         real(c_double), value              :: chemstep            ! minutes
         real(c_double), intent(in)         :: temp(*), pres(*), h2o(*)   ! K, atm, ppm
         real(c_double), intent(in)         :: rj(*)               ! (NCELLS, NPHOTAB) 1/min
         type(c_ptr), value                 :: het, seawater       ! C_LOC(...) or C_NULL_PTR (= 0)
         real(c_double), intent(inout)      :: y(*)                ! (NCELLS, NUMB_MECH_SPC) ppm
         integer(c_int32_t), intent(out)    :: nsteps(*), nfail(*)
         type(cmaq_ros3_fail_t), intent(out) :: fail
      end function cmaq_gear_run

      ! MAP_CHEMISTRY_SPECIES result (RXNS_FUNC_MODULE.F90:1117-1305)
      integer(c_int) function cmaq_ros3_set_cgrid_map(ctx, nspcsd, cgrid_index, convert_conc, &
            species_molwt) bind(c, name='cmaq_ros3_set_cgrid_map')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value                 :: ctx
         integer(c_int32_t), value          :: nspcsd
         integer(c_int32_t), intent(in)     :: cgrid_index(*), convert_conc(*)
         real(c_double), intent(in)         :: species_molwt(*)
      end function cmaq_ros3_set_cgrid_map

      ! gather, CALC_RCONST + RBSOLVER, scatter for one CHEM call (rbdriver.F:444-677)
      integer(c_int) function cmaq_ros3_chem(ctx, grid, chemstep, nacc, nrej, fail) &
            bind(c, name='cmaq_ros3_chem')
         import :: c_int, c_ptr, c_int32_t, c_double, cmaq_ros3_grid_t, cmaq_ros3_fail_t
         type(c_ptr), value                 :: ctx
         type(cmaq_ros3_grid_t), intent(in) :: grid
         real(c_double), value              :: chemstep
         integer(c_int32_t), intent(out)    :: nacc(*), nrej(*)
         type(cmaq_ros3_fail_t), intent(out) :: fail
      end function cmaq_ros3_chem

      ! the failing cell's state (rbsolver.F:532-583); arrays in mechanism order
      integer(c_int) function cmaq_ros3_get_fail_detail(ctx, d, spc_err, y, yp, rj, rki) &
            bind(c, name='cmaq_ros3_get_fail_detail')
         import :: c_int, c_ptr, c_double, cmaq_ros3_fail_detail_t
         type(c_ptr), value                          :: ctx
         type(cmaq_ros3_fail_detail_t), intent(out)  :: d
         real(c_double), intent(out)                 :: spc_err(*), y(*), yp(*), rj(*), rki(*)
      end function cmaq_ros3_get_fail_detail

      integer(c_int) function cmaq_ros3_query(ctx, key, value) bind(c, name='cmaq_ros3_query')
         import :: c_int, c_ptr, c_char, c_int64_t
         type(c_ptr), value                 :: ctx
         character(kind=c_char), intent(in) :: key(*)             ! NUL-terminated
         integer(c_int64_t), intent(out)    :: value
      end function cmaq_ros3_query

      ! INEW2OLD of RBSPARSE, 0-based
      integer(c_int) function cmaq_ros3_species_order(ctx, inew2old) &
            bind(c, name='cmaq_ros3_species_order')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value                 :: ctx
         integer(c_int32_t), intent(out)    :: inew2old(*)
      end function cmaq_ros3_species_order

      type(c_ptr) function cmaq_ros3_last_error(ctx) bind(c, name='cmaq_ros3_last_error')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function cmaq_ros3_last_error

      subroutine cmaq_ros3_free(ctx) bind(c, name='cmaq_ros3_free')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine cmaq_ros3_free

      integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function c_strlen
   end interface

contains

   ! The library's message for the last failure (cmaq_ros3_last_error).
   function cmaq_ros3_error_message(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      p = cmaq_ros3_last_error(ctx)
      if (.not. c_associated(p)) then
         msg = ''
         return
      end if
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [max(n, 1)])
      allocate (character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function cmaq_ros3_error_message

   ! The failing cell's dump as RBSOLVER writes it before M3EXIT
   ! (rbsolver.F:532-583): where and when, MAXERR, then per solver species
   ! (RBSPARSE order, mechanism index and name) its share of ERR, Y and YP;
   ! the cell's met values; RJBLK per photolysis table; the photolysis rate
   ! constants over their RTDAT(1, I) scale.  ncols / nrows (optional) turn
   ! the CGRID-order cell of a CHEM-level call into column, row and layer.
   ! Returns the failure code (0: nothing written).
   integer function cmaq_ros3_write_failure(unit, ctx, jdate, jtime, chemistry_spc, photab, &
         rxlabel, rtdat, iph, ncols, nrows) result(code)
      integer, intent(in)               :: unit, jdate, jtime
      type(c_ptr), intent(in)           :: ctx
      character(len=*), intent(in)      :: chemistry_spc(:), photab(:), rxlabel(:)
      real(c_double), intent(in)        :: rtdat(:, :)             ! RTDAT(3, NRXNS)
      integer(c_int32_t), intent(in)    :: iph(:, :)               ! IPH(NMPHOT, 3)
      integer, intent(in), optional     :: ncols, nrows
      type(cmaq_ros3_fail_detail_t) :: d
      real(c_double), allocatable :: spc_err(:), y(:), yp(:), rj(:), rki(:)
      integer(c_int32_t), allocatable :: inew2old(:)
      integer(c_int) :: rc
      integer :: ns, nr, nj, n, i, j, old
      integer(c_int64_t) :: c
      ns = size(chemistry_spc)
      nr = size(rxlabel)
      nj = size(photab)
      allocate (spc_err(ns), y(ns), yp(ns), rj(max(nj, 1)), rki(nr), inew2old(ns))
      rc = cmaq_ros3_get_fail_detail(ctx, d, spc_err, y, yp, rj, rki)
      code = 0
      if (rc /= CMAQ_ROS3_OK) then
         code = rc
         return
      end if
      code = d%code
      if (d%code == 0) return
      if (d%numb_mech_spc /= ns .or. d%nrxns /= nr .or. d%nphotab /= nj) then
         write (unit, '(a)') ' cmaq_ros3: species, reaction or table count differs from the context'
         return
      end if
      c = d%cell
      write (unit, '(/a, i3, a, i8, a, i7.7, a, i6.6)') ' ROS3 on the GPU: failure (code', d%code, &
         ') in cell', c + 1, '  JDATE = ', jdate, '  JTIME = ', jtime
      if (present(ncols) .and. present(nrows)) then
         write (unit, '(a, 3i6)') '   at column, row, layer:', int(mod(c, int(ncols, c_int64_t))) + 1, &
            int(mod(c / ncols, int(nrows, c_int64_t))) + 1, int(c / (int(ncols, c_int64_t) * nrows)) + 1
      end if
      write (unit, '(a, i5, a, i5)') '   solver species:', ns, '  mechanism species:', ns
      write (unit, '(a, 1pe12.4, a, 1pe12.4, a, 1pe12.4, a, 1pe12.4)') '   MAXERR =', d%maxerr, &
         '  RMS error =', sqrt(d%err / ns), '  DT (min) =', d%dt, '  TNOW (min) =', d%tnow
      write (unit, '(a)') '   index  species           share of ERR        Y (ppm)         YP (ppm)'
      rc = cmaq_ros3_species_order(ctx, inew2old)
      do n = 1, ns
         old = inew2old(n) + 1
         write (unit, '(i8, 2x, a16, 3(1pe17.6))') old, chemistry_spc(old), spc_err(old), y(old), yp(old)
      end do
      write (unit, '(a)') '   TEMP (K)      PRES (atm)    DENS (kg/m3)  H2O (ppm)     SEAWATER'
      write (unit, '(3x, 5(1pe14.5))') d%temp_k, d%pres_atm, d%dens, d%h2o_ppm, d%seawater
      write (unit, '(a)') '   photolysis table        J (1/min)'
      do n = 1, nj
         write (unit, '(i6, 2x, a16, 1pe14.5)') n, photab(n), rj(n)
      end do
      write (unit, '(a)') '   photolysis reaction, table, rate constant / RTDAT(1)'
      do n = 1, size(iph, 1)
         if (iph(n, 3) /= 0) then
            i = iph(n, 1)
            j = iph(n, 2)
            if (rtdat(1, i) > 0.0d0) then
               write (unit, '(3x, a16, 2x, a16, 1pe14.5)') rxlabel(i), photab(j), rki(i) / rtdat(1, i)
            end if
         end if
      end do
   end function cmaq_ros3_write_failure

end module cmaq_ros3_c
